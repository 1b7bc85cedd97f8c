#!/usr/bin/env python3
"""Writes ke_sym_table.hpp: the orbit structure of the 24x24 hexa8 element
stiffness KE0 under the symmetries of the element.

Entry (r, c) = (3a + i, 3b + j) couples component i of local node a with
component j of local node b.  A reflection of axis k moves every node to its
mirror image and flips displacement component k; an axis permutation moves
nodes and components alike; KE0 is symmetric (transpose).  Entries in one
orbit are equal up to the sign the group element carries, so KE0 is fixed by
one value per orbit:
    KE0[r][c] = sign * v[orbit]      code[r*24 + c] = sign * (orbit + 1)
BOX  group: reflections x transpose (order 16), any box element.
CUBE group: reflections x axis permutations x transpose (order 96), cubic
elements of an isotropic material.
Local node order: bottom face counter-clockwise 0..3, top face 4..7.

Usage: python gen_ke_sym.py > ke_sym_table.hpp
"""
import itertools


def lnode_off(a):
    x = 1 if (a & 3) in (1, 2) else 0
    y = 1 if (a & 3) >= 2 else 0
    return (x, y, a >> 2)


def lnode(o):
    x, y, z = o
    return (0 if y == 0 else 3) + 4 * z if x == 0 else (1 if y == 0 else 2) + 4 * z


def act(perm, mask, a, i):
    """Image of (node a, component i) and the sign of the component."""
    o = lnode_off(a)
    r = [o[k] ^ ((mask >> k) & 1) for k in range(3)]
    p = [0, 0, 0]
    for k in range(3):
        p[perm[k]] = r[k]
    sign = -1 if (mask >> i) & 1 else 1
    return lnode(tuple(p)), perm[i], sign


def orbits(perms):
    rep = {}
    for r in range(24):
        for c in range(24):
            a, i, b, j = r // 3, r % 3, c // 3, c % 3
            best = None
            for perm in perms:
                for mask in range(8):
                    a2, i2, s1 = act(perm, mask, a, i)
                    b2, j2, s2 = act(perm, mask, b, j)
                    for rr, cc in (((3 * a2 + i2), (3 * b2 + j2)), ((3 * b2 + j2), (3 * a2 + i2))):
                        # image entry = s1*s2 * entry  =>  entry = s1*s2 * image
                        cand = (rr * 24 + cc, s1 * s2)
                        if best is None or cand[0] < best[0]:
                            best = cand
            rep[r * 24 + c] = best
    reps = sorted({v[0] for v in rep.values()})
    index = {q: n for n, q in enumerate(reps)}
    codes = [rep[e][1] * (index[rep[e][0]] + 1) for e in range(576)]
    return codes, reps


def emit(name, codes, reps):
    out = [f"// {len(reps)} orbits; representative entries (row*24+col):",
           "// " + ", ".join(map(str, reps)),
           f"constexpr int {name}_N = {len(reps)};",
           f"constexpr int {name}_REP[{len(reps)}] = {{{', '.join(map(str, reps))}}};",
           f"constexpr signed char {name}[576] = {{"]
    for r in range(24):
        out.append("    " + ", ".join(f"{v:3d}" for v in codes[24 * r:24 * r + 24]) + ",")
    out.append("};")
    return "\n".join(out)


def main():
    box = orbits([(0, 1, 2)])
    cube = orbits(list(itertools.permutations(range(3))))
    print("// ke_sym_table.hpp -- generated by gen_ke_sym.py; do not edit.")
    print("// Orbit codes of the hexa8 KE0 entries: KE0[r][c] = sign(code) * v[|code| - 1].")
    print("#pragma once\n\nnamespace icarat_hip {\n")
    print(emit("KSYM_BOX", *box))
    print()
    print(emit("KSYM_CUBE", *cube))
    print("\n}  // namespace icarat_hip")


if __name__ == "__main__":
    main()
