"""CPU restatement of the beta -> alpha driver's output conversions (test
infrastructure only: tests/ check libetd's GPU conversions against it).

field2b     Dtype2B_E, source_beta_to_alpha/src/anisotropic_hip.c:3920-3931:
            (Stype)((E + 0.1) * 30000) on every cell of the extended field.
            The double -> unsigned short cast runs on the GPU there; the GPU
            converts double -> uint32 truncating toward zero and saturating
            (negative and NaN -> 0, >= 2^32 -> 2^32 - 1) and keeps the low 16
            bits.  Pinned by known answers at -0.1, 0, 0.9, 2.0 (in range),
            2.5 (wraps), -1 (saturates) and NaN.
downsample  skip_copyin_new_2, anisotropic_hip.c:3897-3918: the (N/2)^3 means
            of 2x2x2 interior blocks, x fastest, as the reference adds the 8
            cells (left to right, in FP64) and multiplies by 1/8.  Pinned
            against numpy's block mean away from rounding ties and by
            hand-made blocks.
detector    the binarisation of anisotropic_hip.c:4264-4278 (mean > 0.1 ->
            1.0f) and float2bin, 4400-4427 (byte t = cells 8t .. 8t+7, bit e
            = cell 8t + e).  Pinned by numpy.packbits(bitorder="little") and
            single-bit known answers.
"""
import numpy as np


def field2b(u):
    x = (np.asarray(u, dtype=np.float64) + 0.1) * 30000.0
    w = np.where(np.isnan(x), 0.0, np.clip(np.trunc(x), 0.0, 4294967295.0)).astype(np.uint64)
    return (w & 0xFFFF).astype(np.uint16)


def downsample(u, g=2):
    u = np.asarray(u, dtype=np.float64)
    n = u.shape[0]
    N = n - 2 * g
    d = N // 2
    out = np.empty((d, d, d))
    for k in range(d):
        for j in range(d):
            z, y = 2 * k + g, 2 * j + g
            x = 2 * np.arange(d) + g
            s = (u[z, y, x] + u[z, y, x + 1] + u[z, y + 1, x] + u[z, y + 1, x + 1] + u[z + 1, y, x] +
                 u[z + 1, y, x + 1] + u[z + 1, y + 1, x] + u[z + 1, y + 1, x + 1])
            out[k, j] = (1.0 / 8.0) * s
    return out


def detector(u, g=2):
    """(bits, binary): bits (N/2)^3 / 8 uint8, binary (N/2)^3 float32 (x fastest)."""
    m = downsample(u, g).reshape(-1)
    binary = np.where(m > 0.1, np.float32(1.0), np.float32(0.0)).astype(np.float32)
    on = (m > 0.1).reshape(-1, 8)
    bits = np.zeros(on.shape[0], np.uint32)
    for e in range(8):
        bits |= on[:, e].astype(np.uint32) << e
    return bits.astype(np.uint8), binary
