"""Lattice velocity tables in the reference ordering (vars_module.f90:319-338;
the first 19 entries are the D3Q19 set of vars_module.f90:666-684)."""
EX = (0, 1, -1, 0, 0, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, -1, 1, 1, -1)
EY = (0, 0, 0, 1, -1, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, -1, 1, -1, 1)
EZ = (0, 0, 0, 0, 0, 1, -1, 0, 0, 0, 0, 1, -1, -1, 1, 1, -1, 1, -1, 1, -1, 1, -1, 1, -1, -1, 1)
OPP = (0, 2, 1, 4, 3, 6, 5, 8, 7, 10, 9, 12, 11, 14, 13, 16, 15, 18, 17, 20, 19, 22, 21, 24, 23, 26, 25)


def weights(lattice: int):
    out = []
    n = 27 if lattice == 27 else 19
    for l in range(n):
        d2 = EX[l] ** 2 + EY[l] ** 2 + EZ[l] ** 2
        if lattice == 27:
            out.append((8 / 27, 2 / 27, 1 / 54, 1 / 216)[d2])
        else:
            out.append((1 / 3, 1 / 18, 1 / 36)[d2])
    return tuple(out)
