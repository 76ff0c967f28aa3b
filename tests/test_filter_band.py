"""CPU checks of the MFMA filter's error band (hj_filter_band, host only).

The filter's arithmetic is emulated exactly as the fragment builder and the
MFMA chain perform it: operands y = (x - c) * scale rounded FP64 -> FP32 ->
FP16 / BF16 (round to nearest even), squared norms summed in FP64 and split
into FP16 / BF16 parts against -2^15 (in the last slots of each K chunk),
products exact in FP32, and the 16-slot k-steps accumulated in FP32 in
several orders (the MFMA's internal order is not specified).  Pairs are built adversarially: coordinates of magnitude near
the radius bound, each one placed almost half an ulp from a rounding boundary
so that every dimension's rounding error pushes the pair's distance the same
way, at true distances swept across both edges of the band.  No pair that the
thresholds call certainly within eps (acc >= t_in) or certainly outside
(acc < t_out) may disagree with the fixed-order FP64 predicate.
"""
import numpy as np
import pytest

import hilbertjoin_amd as hj
import oracle

NORM_UNIT = 32768.0


def round_op(x, bf16):
    f = np.asarray(x, dtype=np.float64).astype(np.float32)
    if not bf16:
        return f.astype(np.float16).astype(np.float64)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def ulp_op(y, bf16):
    """Spacing of the operand format at |y| (normal range)."""
    e = np.floor(np.log2(np.maximum(np.abs(y), 2.0 ** -14)))
    return 2.0 ** (e - (7 if bf16 else 10))


def slots(y, fi, d, row, bf16):
    """K slots of one point in the row or column role, all k-steps."""
    waves = 1 if d <= 32 else 4
    split, parts = fi["split"], fi["parts"]
    # norms of the two chunks: per wave group k = g, g + waves, ... then summed
    # (y has at most 11 significant bits, so y*y is exact and fma(y, y, s)
    # equals y*y + s)
    norms = []
    for lo, hi in ((0, split), (split, d)):
        grp = [0.0] * waves
        for k in range(d):
            if lo <= k < hi:
                grp[k % waves] = float(y[k]) * float(y[k]) + grp[k % waves]
        v = 0.0
        for w in range(waves):
            v += grp[w]
        norms.append(v * (0.5 / NORM_UNIT))
    out = []
    for chunk, (lo, hi) in enumerate(((0, split), (split, d))):
        if lo >= hi:
            continue
        v = norms[chunk]
        pr = []
        for _ in range(parts):
            p = float(round_op(v, bf16))
            pr.append(p)
            v -= p
        data = list(y[lo:hi])
        if row:
            extra = pr + [-NORM_UNIT] * parts
        else:
            extra = [-NORM_UNIT] * parts + pr
        # the chunk's norm slots are the last 2 * parts slots of its k-steps
        nslots = -(-(len(data) + len(extra)) // 16) * 16
        s = data + [0.0] * (nslots - len(data) - len(extra)) + extra
        out.extend(s)
    assert len(out) == 16 * fi["steps"]
    return np.array(out)


def accumulate(pa, pb, order):
    """FP32 accumulation of the exact products, k-step by k-step."""
    prod = pa * pb  # exact: operands have at most 11 significant bits
    steps = prod.reshape(-1, 16)
    acc = np.float32(0.0)
    for st in steps:
        if order == "seq":
            for v in st:
                acc = np.float32(acc + np.float32(v))
        elif order == "rev":
            for v in st[::-1]:
                acc = np.float32(acc + np.float32(v))
        elif order == "tree":
            t = [np.float32(v) for v in st]
            while len(t) > 1:
                t = [np.float32(t[i] + t[i + 1]) for i in range(0, len(t), 2)]
            acc = np.float32(acc + t[0])
        elif order == "once":  # the k-step's exact sum, rounded once
            acc = np.float32(float(acc) + float(np.sum(st)))
    return float(acc)


def adversarial_pairs(rng, d, eps, radius, fi, bf16, closer, taus):
    s = fi["scale"]
    M = s * radius
    pairs = []
    for tau in taus:
        w = rng.normal(size=d)
        w /= np.linalg.norm(w)
        a0 = rng.normal(size=d)
        # as far out as the radius bound allows for both points (large
        # magnitudes maximize the rounding and accumulation errors)
        room = M - s * eps * (1.0 + abs(tau)) - 2.0 * np.sqrt(d) * ulp_op(M, bf16)
        a0 *= min(0.9 * M, room) / np.linalg.norm(a0)
        b0 = a0 + s * eps * (1.0 + tau) * w
        # place every coordinate ~half an ulp from a rounding boundary so that
        # the rounded difference shrinks (closer) or grows in every dimension
        sgn = np.sign(w) * (1.0 if closer else -1.0)
        ya, yb = np.empty(d), np.empty(d)
        for k in range(d):
            for y0, dst, want in ((a0[k], ya, sgn[k]), (b0[k], yb, -sgn[k])):
                u = ulp_op(y0, bf16)
                m = np.floor(y0 / u)
                # m u + (0.5 -+ 0.001) u rounds to m u (error < 0) or (m+1) u (> 0)
                dst[k] = (m + 0.5 + (0.001 if want > 0 else -0.001)) * u
        pairs.append((ya / s, yb / s))  # exact: s is a power of two
    return pairs


CASES = [("f16", 4, 0.3, 1.0), ("f16", 16, 0.06, 2.0), ("f16", 28, 0.03, 2.65),
         ("f16", 64, 0.2, 4.0), ("bf16", 4, 0.3, 1.0), ("bf16", 16, 0.06, 2.0),
         ("bf16", 28, 0.2, 2.65), ("bf16", 64, 0.2, 4.0)]


@pytest.mark.parametrize("kernel,d,eps,radius", CASES)
def test_thresholds_never_misclassify_adversarial_pairs(kernel, d, eps, radius):
    bf16 = kernel == "bf16"
    fi = hj.filter_band(kernel, d, eps, radius)
    s = fi["scale"]
    assert fi["t_out"] < fi["t_in"] or fi["t_in"] > 0
    # band edges as relative distances: acc = -(s D)^2 / 2 at the thresholds
    d_out = np.sqrt(-2.0 * fi["t_out"]) / (s * eps) - 1.0
    d_in = 1.0 - np.sqrt(max(-2.0 * fi["t_in"], 0.0)) / (s * eps)
    w = max(d_out, d_in)
    taus = np.concatenate([np.linspace(-3 * w, 3 * w, 61), np.linspace(-1.2 * w, 1.2 * w, 61)])
    rng = np.random.default_rng(d * 7 + (1 if bf16 else 0))
    decided_in = decided_out = 0
    for closer in (True, False):
        for xa, xb in adversarial_pairs(rng, d, eps, radius, fi, bf16, closer, taus):
            assert np.linalg.norm(xa) <= radius and np.linalg.norm(xb) <= radius
            truth = oracle.pair_within(xa, xb, eps)
            ya, yb = round_op(xa * s, bf16), round_op(xb * s, bf16)
            ra = slots(ya, fi, d, True, bf16)
            cb = slots(yb, fi, d, False, bf16)
            for order in ("seq", "rev", "tree", "once"):
                acc = accumulate(ra, cb, order)
                if acc >= fi["t_in"]:
                    assert truth, (order, acc, fi)
                    decided_in += 1
                if acc < fi["t_out"]:
                    assert not truth, (order, acc, fi)
                    decided_out += 1
                # the first K chunk alone (the filter's early exit) is a
                # partial distance: certainly-out there must be out too
                k1 = 16 * fi["steps1"]
                if fi["steps1"] < fi["steps"]:
                    acc1 = accumulate(ra[:k1], cb[:k1], order)
                    if acc1 < fi["t_out"]:
                        assert not truth
    # the band is not vacuous: pairs a few band widths away are decided
    assert decided_out > 0
    if fi["t_in"] < 0:
        assert decided_in > 0


def test_band_shrinks_with_radius_and_precision():
    # the band's relative width grows with radius / eps and with the operand
    # unit roundoff (BF16 keeps 8 bits, FP16 11)
    def rel(kernel, radius):
        fi = hj.filter_band(kernel, 28, 0.05, radius)
        return np.sqrt(-2.0 * fi["t_out"]) / (fi["scale"] * 0.05) - 1.0
    assert rel("f16", 0.5) < rel("f16", 2.0) < rel("f16", 8.0)
    assert rel("f16", 2.0) < rel("bf16", 2.0)
    with pytest.raises(hj.HJError):
        hj.filter_band("fp64", 28, 0.05, 1.0)
    with pytest.raises(hj.HJError):
        hj.filter_band("f16", 28, -1.0, 1.0)
