#!/usr/bin/env python3
"""Generate su2_mrg32k3a_jump.h: the MRG32k3a transition matrices raised to
2^(76+i), i = 0..NJUMP-1, for both components.

MRG32k3a (L'Ecuyer, "Good parameters and implementations for combined
multiple recursive random number generators", Oper. Res. 47 (1999)) has two
order-3 recurrences
    x1[n] = (1403580 x1[n-2] - 810728 x1[n-3]) mod m1,  m1 = 2^32 - 209
    x2[n] = (527612 x2[n-1] - 1370589 x2[n-3]) mod m2,  m2 = 2^32 - 22853
With the state vector s = (x[n-3], x[n-2], x[n-1]) one step is s <- A s.
A subsequence is 2^76 steps long (the RngStreams layout of L'Ecuyer, Simard,
Chen and Kelton, Oper. Res. 50 (2002)); jumping to subsequence k multiplies
the state by A^(2^76 k), i.e. by the product of the tabulated
A^(2^(76+i)) over the set bits i of k.  Exact Python integers, so the table
is the mathematical matrix power.

Run:  python3 gen_mrg32k3a_jump.py > su2_mrg32k3a_jump.h
"""

M1 = 4294967087
M2 = 4294944443
A1 = [[0, 1, 0], [0, 0, 1], [(-810728) % M1, 1403580, 0]]
A2 = [[0, 1, 0], [0, 0, 1], [(-1370589) % M2, 0, 527612]]
NJUMP = 32  # subsequences < 2^32 (one per half-site)


def matmul(a, b, m):
    return [[sum(a[i][k] * b[k][j] for k in range(3)) % m for j in range(3)] for i in range(3)]


def matpow2(a, e, m):
    """a^(2^e) by e squarings."""
    r = a
    for _ in range(e):
        r = matmul(r, r, m)
    return r


def table(a, m):
    out = []
    r = matpow2(a, 76, m)
    for _ in range(NJUMP):
        out.append(r)
        r = matmul(r, r, m)
    return out


def emit(name, tab):
    lines = [f"static const uint32_t {name}[{NJUMP} * 9] = {{"]
    for mat in tab:
        flat = [v for row in mat for v in row]
        lines.append("    " + ", ".join(f"{v}u" for v in flat) + ",")
    lines.append("};")
    return "\n".join(lines)


def main():
    print("/* Generated by gen_mrg32k3a_jump.py - do not edit.")
    print(" * MRG32k3a transition matrices A^(2^(76+i)), i = 0..%d, row-major," % (NJUMP - 1))
    print(" * for component 1 (mod 4294967087) and component 2 (mod 4294944443). */")
    print("#ifndef SU2_MRG32K3A_JUMP_H")
    print("#define SU2_MRG32K3A_JUMP_H")
    print("#include <stdint.h>")
    print(f"#define SU2_MRG_NJUMP {NJUMP}")
    print(emit("su2_mrg_jump1_host", table(A1, M1)))
    print(emit("su2_mrg_jump2_host", table(A2, M2)))
    print("#endif")


if __name__ == "__main__":
    main()
