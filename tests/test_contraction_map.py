"""The matrix-core contractions' mappings (contract_mfma_bf16, contract_lines_bf16 and
contract_lines_f32 in 3ddg_amd/csrc/dg3d_kernels.hip), emulated on the CPU for every stage shape
the library builds.

The emulation restates the kernel's index arithmetic: the column -> line map (column_line), the
line descriptors, the aligned diagonal blocks of the 16x16x16 tiles (one every 8 rows), the lane
layouts of v_mfma_f32_16x16x16bf16_1k, v_mfma_f32_4x4x4bf16_1k and v_mfma_f32_4x4x1f32, and the
round-robin of tiles and line groups over waves. It checks what the kernel relies on without a
bounds check of its own: every LDS read and write stays inside its array, every output line is
written exactly once, and the result equals the direct per-node contraction
out[i] = sum_m A[i][m] line[m] along each axis, where a P/rho0 line enters shifted by its middle
value (the FP32 stage's shifted pressure lines).
"""
import numpy as np
import pytest

# stage shapes: elements per workgroup (both autotuner shapes) per nodes-per-line n
SHAPES = {2: (32, 16), 3: (8, 16), 4: (4, 8), 5: (2, 4), 6: (2, 1), 7: (2, 1), 8: (1, 2)}


def elem_stride(n):
    """Node slots per element (128-byte lines of doubles), as dg3d_device.hpp's elem_stride."""
    return n ** 3 if (n == 2 or n ** 3 % 32 == 0) else (n ** 3 + 31) // 32 * 32


def column_line(n, col):
    """(element, axis, field, face node) of a column -> first node offsets in sF and sOut, and
    the node stride along the line."""
    NP, NF2 = n ** 3, n * n
    le, rr = divmod(col, 6 * NF2)
    af, fn = divmod(rr, NF2)
    a, f = af >> 1, af & 1
    b, c = divmod(fn, n)
    base = b * n + c if a == 0 else (b * NF2 + c if a == 1 else b * NF2 + c * n)
    ls = NF2 if a == 0 else (n if a == 1 else 1)
    return le * 4 * NP + (0 if f == 0 else 1 + a) * NP + base, (le * 6 + af) * NP + base, ls


def emulate(n, epb, A, sF):
    """The 16x16x16 split-BF16 tiles (contract_mfma_bf16) in exact arithmetic."""
    NP, NF2 = n ** 3, n * n
    NT = (epb * elem_stride(n) + 63) // 64 * 64
    W = NT // 64
    BS = 8
    P = 16 // BS
    C = epb * 6 * NF2
    TC = 16 * P
    T = (C + TC - 1) // TC
    sOut = np.full(epb * 6 * NP, np.nan)
    writes = np.zeros(epb * 6 * NP, int)

    def line(col):
        ok = col < C
        c = col if ok else 0
        fo, oo, ls = column_line(n, c)
        assert 0 <= fo < sF.size
        pressure = ((c % (6 * NF2)) // NF2) % 2 == 0
        assert 0 <= fo + (n // 2) * ls < sF.size
        return ok, fo, oo, ls, (sF[fo + (n // 2) * ls] if pressure else 0.0)

    def load(ln, m):
        ok, fo, _, ls, sh = ln
        if ok and m < n:
            i = fo + m * ls
            assert 0 <= i < sF.size
            return sF[i] - sh
        return 0.0

    for wave in range(W):
        for t in range(wave, T, W):
            Aop = np.zeros((16, 16))
            Bop = np.zeros((16, 16))
            out = []
            for lane in range(64):
                j, g = lane & 15, lane >> 4
                bj, mj = divmod(j, BS)
                bd = (4 * g) // BS
                md = 4 * g - bd * BS
                # K rows 4g..4g+3 of the lane, one block
                ln = line(t * TC + 16 * bd + j)
                for q in range(4):
                    k, mk = 4 * g + q, md + q
                    Aop[j, k] = A[mj, mk] if (bj == bd and mj < n and mk < n) else 0.0
                    Bop[k, j] = load(ln, mk)
                out.append((j, [(4 * g + r, ln, md + r) for r in range(4)]))
            D = Aop @ Bop
            for j, rows in out:
                for row, (ok, _, oo, ls, _sh), m in rows:
                    if ok and m < n:
                        i = oo + m * ls
                        assert 0 <= i < sOut.size
                        sOut[i] = D[row, j]
                        writes[i] += 1
    return sOut, writes


def direct(n, epb, A, sF):
    NP, NF2 = n ** 3, n * n
    ref = np.zeros(epb * 6 * NP)
    for le in range(epb):
        for node in range(NP):
            idx = (node // NF2, (node // n) % n, node % n)
            for a in range(3):
                ls = NF2 if a == 0 else (n if a == 1 else 1)
                base = node - idx[a] * ls
                F = le * 4 * NP
                mid = sF[F + base + (n // 2) * ls]
                ref[(le * 6 + 2 * a) * NP + node] = sum(A[idx[a], m] * (sF[F + base + m * ls] - mid)
                                                        for m in range(n))
                ref[(le * 6 + 2 * a + 1) * NP + node] = sum(
                    A[idx[a], m] * sF[F + (1 + a) * NP + base + m * ls] for m in range(n))
    return ref


@pytest.mark.parametrize("n", [k for k in sorted(SHAPES) if k > 4])
def test_tile_map_covers_every_line_once_in_bounds(n):
    """The 16x16x16 split-BF16 tiles run at 8 nodes per line; their map holds from 5 up."""
    rng = np.random.default_rng(n)
    A = rng.standard_normal((n, n))
    for epb in SHAPES[n]:
        sF = rng.standard_normal(epb * 4 * n ** 3)
        ref = direct(n, epb, A, sF)
        got, writes = emulate(n, epb, A, sF)
        assert (writes == 1).all(), (n, epb, writes.min(), writes.max())
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def descriptor(n, epb, col):
    """The split-BF16 line descriptor (LineTable in dg3d_kernels.hip): bits 0-11 the line's first
    node in sF, 12-24 in sOut, 25-26 the axis, 27 a P/rho0 line, 31 a line that exists; 0 for the
    padding lines of the last group of 64."""
    NF2 = n * n
    if col >= epb * 6 * NF2:
        return 0
    fo, oo, _ = column_line(n, col)
    a = (col % (6 * NF2)) // NF2 >> 1
    pressure = ((col % (6 * NF2)) // NF2) % 2 == 0
    assert fo < 4096 and oo < 8192
    return fo | oo << 12 | a << 25 | int(pressure) << 27 | 1 << 31


def mfma_4x4x4(A_lanes, B_lanes):
    """v_mfma_f32_4x4x4bf16_1k as probed on the device: 16 blocks of 4 lanes; lane 4b + i holds
    row i of block b's A (4 K values), lane 4b + j column j of block b's B, and register r of lane
    4b + j receives row r, column j of block b's product."""
    D = np.zeros((64, 4))
    for b in range(16):
        Ab = A_lanes[4 * b:4 * b + 4]          # [i][k]
        Bb = B_lanes[4 * b:4 * b + 4].T        # [k][j]
        D[4 * b:4 * b + 4] = (Ab @ Bb).T       # lane j, register r
    return D


def emulate_lines(n, epb, A, sF):
    """The line-per-lane split-BF16 contraction (contract_lines_bf16) in exact arithmetic: each
    lane decodes its line from its descriptor, loads it (P/rho0 lines shifted by their middle
    value), and takes its rows from the 16-block 4x4x4 products of the pieces of A."""
    NP, NF2 = n ** 3, n * n
    NT = (epb * elem_stride(n) + 63) // 64 * 64
    W = NT // 64
    C = epb * 6 * NF2
    G = (C + 63) // 64
    RB = KB = 2 if n > 4 else 1
    Ap = np.zeros((4 * RB, 4 * KB))
    Ap[:n, :n] = A
    sOut = np.full(epb * 6 * NP, np.nan)
    writes = np.zeros(epb * 6 * NP, int)
    for wave in range(W):
        for g in range(wave, G, W):
            x = np.zeros((64, 4 * KB))
            dec = []
            for lane in range(64):
                d = descriptor(n, epb, g * 64 + lane)
                ok = d >> 31 & 1
                fo = d & 0xFFF if ok else 0
                oo, ax, pr = d >> 12 & 0x1FFF, d >> 25 & 3, d >> 27 & 1
                ls = NF2 if ax == 0 else (n if ax == 1 else 1)
                for m in range(n):
                    assert 0 <= fo + m * ls < sF.size
                    x[lane, m] = sF[fo + m * ls]
                sh = x[lane, n // 2] if pr else 0.0
                x[lane, :n] = (x[lane, :n] - sh) if ok else 0.0
                dec.append((ok, oo, ls))
            for R in range(RB):
                out = np.zeros((64, 4))
                for K in range(KB):
                    A_lanes = np.array([Ap[4 * R + (lane & 3), 4 * K:4 * K + 4] for lane in range(64)])
                    out += mfma_4x4x4(A_lanes, x[:, 4 * K:4 * K + 4])
                for lane, (ok, oo, ls) in enumerate(dec):
                    for r in range(4):
                        row = 4 * R + r
                        if ok and row < n:
                            i = oo + row * ls
                            assert 0 <= i < sOut.size
                            sOut[i] = out[lane, r]
                            writes[i] += 1
    return sOut, writes


@pytest.mark.parametrize("n", [k for k in sorted(SHAPES) if k < 8])
def test_line_per_lane_map_covers_every_line_once_in_bounds(n):
    """At 2..7 nodes per line the split-BF16 contraction takes one line per lane: every line is
    written exactly once, every LDS index is in bounds, and the result equals the direct per-node
    contraction; the pieces of A sit on a grid of 4x4 blocks (one for n <= 4, four above)."""
    rng = np.random.default_rng(100 + n)
    A = rng.standard_normal((n, n))
    for epb in SHAPES[n]:
        sF = rng.standard_normal(epb * 4 * n ** 3)
        ref = direct(n, epb, A, sF)
        got, writes = emulate_lines(n, epb, A, sF)
        assert (writes == 1).all(), (n, epb, writes.min(), writes.max())
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())


def mfma_4x4x1(a_lanes, b_lanes):
    """v_mfma_f32_4x4x1f32 as probed on the device: 16 blocks of 4 lanes; lane 4b + i holds row i
    of block b's A (one K value), lane 4b + j column j of its B, and register r of lane 4b + j
    receives row r, column j of block b's outer product."""
    D = np.zeros((64, 4))
    for b in range(16):
        D[4 * b:4 * b + 4] = np.outer(b_lanes[4 * b:4 * b + 4], a_lanes[4 * b:4 * b + 4])
    return D


def emulate_lines_f32(n, epb, A, sF):
    """The line-per-lane FP32 contraction (contract_lines_f32): the same line descriptors as the
    split-BF16 one; each lane feeds row 4R + (lane & 3) of A's column k and its own line's node k,
    one product per node and row block."""
    NP, NF2 = n ** 3, n * n
    NT = (epb * elem_stride(n) + 63) // 64 * 64
    W = NT // 64
    C = epb * 6 * NF2
    G = (C + 63) // 64
    RB = (n + 3) // 4
    sOut = np.full(epb * 6 * NP, np.nan)
    writes = np.zeros(epb * 6 * NP, int)
    for wave in range(W):
        for g in range(wave, G, W):
            x = np.zeros((64, n))
            dec = []
            for lane in range(64):
                d = descriptor(n, epb, g * 64 + lane)
                ok = d >> 31 & 1
                fo = d & 0xFFF if ok else 0
                oo, ax, pr = d >> 12 & 0x1FFF, d >> 25 & 3, d >> 27 & 1
                ls = NF2 if ax == 0 else (n if ax == 1 else 1)
                for m in range(n):
                    assert 0 <= fo + m * ls < sF.size
                    x[lane, m] = sF[fo + m * ls]
                sh = x[lane, n // 2] if pr else 0.0
                x[lane] = (x[lane] - sh) if ok else 0.0
                dec.append((ok, oo, ls))
            for R in range(RB):
                out = np.zeros((64, 4))
                for k in range(n):
                    a_lanes = np.array([A[4 * R + (lane & 3), k] if 4 * R + (lane & 3) < n else 0.0
                                        for lane in range(64)])
                    out += mfma_4x4x1(a_lanes, x[:, k])
                for lane, (ok, oo, ls) in enumerate(dec):
                    for r in range(4):
                        row = 4 * R + r
                        if ok and row < n:
                            i = oo + row * ls
                            assert 0 <= i < sOut.size
                            sOut[i] = out[lane, r]
                            writes[i] += 1
    return sOut, writes


@pytest.mark.parametrize("n", sorted(SHAPES))
def test_fp32_line_per_lane_map_covers_every_line_once_in_bounds(n):
    """The FP32 matrix-core contraction takes one line per lane at every order: every line is
    written exactly once, every LDS index is in bounds, and the result equals the direct per-node
    contraction."""
    rng = np.random.default_rng(200 + n)
    A = rng.standard_normal((n, n))
    for epb in SHAPES[n]:
        sF = rng.standard_normal(epb * 4 * n ** 3)
        ref = direct(n, epb, A, sF)
        got, writes = emulate_lines_f32(n, epb, A, sF)
        assert (writes == 1).all(), (n, epb, writes.min(), writes.max())
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * np.abs(ref).max())
