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


# ---- tile frames -------------------------------------------------------------------
# A tile pair with its own frames: operands y~ = round(z - anchor), anchors on
# the lattice, the norm slots corrected by r = y~_a . delta and c = y~_b .
# delta (FP32 dot products of 2-slot pairs per lane half, added over the
# chunk's k-steps, the two lane halves summed last) and re-split in FP32,
# thresholds per pair from hj_filter_pair_band.  Emulated as the kernels do it,
# with several FP32 rounding orders where the hardware's is not specified.

def round_f32(x):
    return float(np.float32(x))


def to_op(x, bf16):
    return float(round_op(np.array([x]), bf16)[0])


def dot2(acc, p0, p1, mode):
    """One 2-slot dot product step (products exact in FP32)."""
    if mode == "fused":
        return round_f32(acc + p0 + p1)
    if mode == "seq":
        return round_f32(round_f32(acc + p0) + p1)
    return round_f32(round_f32(p0 + p1) + acc)  # "pairfirst"


def chunk_dot(op_slots, dl_slots, steps, mode):
    """Per lane half: 4 dot2 steps per k-step, over the chunk's k-steps; then
    the two halves' sum."""
    halves = [0.0, 0.0]
    for t in range(steps):
        for h in range(2):
            base = 16 * t + 8 * h
            for w in range(4):
                a0, a1 = op_slots[base + 2 * w], op_slots[base + 2 * w + 1]
                d0, d1 = dl_slots[base + 2 * w], dl_slots[base + 2 * w + 1]
                halves[h] = dot2(halves[h], a0 * d0, a1 * d1, mode)
    return round_f32(halves[0] + halves[1])


def chunk_dd(dl_slots, steps):
    halves = [0.0, 0.0]
    for t in range(steps):
        for h in range(2):
            for j in range(8):
                v = dl_slots[16 * t + 8 * h + j]
                halves[h] = round_f32(halves[h] + v * v)  # an FMA: one rounding
    return round_f32(halves[0] + halves[1])


def resplit(parts_vals, corr, parts, bf16):
    """adjust_norm: the parts' FP32 sum (in order; a subnormal low part can
    make it inexact) plus corr / 2^15 in one rounding, split again."""
    v = parts_vals[0]
    for p in parts_vals[1:]:
        v = round_f32(v + p)
    v = round_f32(v + corr / 32768.0)
    q0 = to_op(v, bf16)
    r1 = round_f32(v - q0)
    if parts == 2:
        return [q0, to_op(r1, bf16)]
    q1 = to_op(r1, bf16)
    return [q0, q1, to_op(round_f32(r1 - q1), bf16)]


def corrected(sl, chunks, corr, parts, row, bf16):
    """Slots of one operand with each chunk's norm parts corrected."""
    out = list(sl)
    for (t0, t1), cv in zip(chunks, corr):
        n0 = 16 * t1 - 2 * parts + (0 if row else parts)  # the role's parts
        out[n0:n0 + parts] = resplit(out[n0:n0 + parts], cv, parts, bf16)
    return np.array(out)


def delta_slots(dl, fi, d):
    """delta per K slot in the fragment layout (0 off the data slots)."""
    split, parts = fi["split"], fi["parts"]
    out = []
    for lo, hi in ((0, split), (split, d)):
        if lo >= hi:
            continue
        data = list(dl[lo:hi])
        nslots = -(-(len(data) + 2 * parts) // 16) * 16
        out.extend(data + [0.0] * (nslots - len(data)))
    return np.array(out)


def adversarial_frame_pairs(rng, d, s, eps, q, rho, taus, bf16, closer):
    """(y~_a, y~_b, delta, x_a, x_b): z_a = a_I + y_a, z_b = a_J + y_b with |y| ~ rho,
    |z_a - z_b| = s eps (1 + tau), every operand coordinate placed ~half an ulp
    from a rounding boundary so that the rounding moves the pair's distance the
    same way in every dimension."""
    out = []
    for tau in taus:
        w = rng.normal(size=d)
        w /= np.linalg.norm(w)
        ya0 = rng.normal(size=d)
        ya0 *= rho / np.linalg.norm(ya0)
        yb0 = rng.normal(size=d)
        yb0 *= rho / np.linalg.norm(yb0)
        a_i = q * np.round(rng.normal(size=d) * 4)
        # anchors' difference on the lattice, chosen so that z_b lands near a_J + yb0
        delta = q * np.round((ya0 + s * eps * (1.0 + tau) * w - yb0) / q)
        a_j = a_i + delta
        za0 = a_i + ya0
        zb0 = za0 + s * eps * (1.0 + tau) * w
        sgn = np.sign(w) * (1.0 if closer else -1.0)
        ya, yb = np.empty(d), np.empty(d)
        for k in range(d):
            for y0, dst, want in ((za0[k] - a_i[k], ya, sgn[k]), (zb0[k] - a_j[k], yb, -sgn[k])):
                u = ulp_op(y0, bf16)
                m = np.floor(y0 / u)
                dst[k] = (m + 0.5 + (0.001 if want > 0 else -0.001)) * u
        za, zb = a_i + ya, a_j + yb  # exact: lattice and ulp multiples
        out.append((round_op(ya, bf16), round_op(yb, bf16), delta, za / s, zb / s))
    return out


FRAME_CASES = [("f16", 28, 0.05, 2.65, 2.0), ("f16", 16, 0.06, 2.0, 6.0),
               ("f16", 64, 0.2, 4.0, 1.5), ("bf16", 28, 0.05, 2.65, 2.0),
               ("bf16", 16, 0.06, 2.0, 6.0), ("bf16", 64, 0.2, 4.0, 1.5)]


@pytest.mark.parametrize("kernel,d,eps,radius,rho_eps", FRAME_CASES)
def test_frame_thresholds_never_misclassify_adversarial_pairs(kernel, d, eps, radius, rho_eps):
    bf16 = kernel == "bf16"
    fi = hj.filter_band(kernel, d, eps, radius)
    s, parts, steps1, steps = fi["scale"], fi["parts"], fi["steps1"], fi["steps"]
    assert steps1 == 1  # tile frames need a one-k-step first chunk
    amax = 128 if bf16 else 1024
    q = 2.0 ** np.ceil(np.log2(s * radius / amax))
    rho = rho_eps * s * eps  # the tiles' radius in their own frames
    chunks = [(0, steps1)] + ([(steps1, steps)] if steps > steps1 else [])
    rng = np.random.default_rng(d * 11 + (1 if bf16 else 0))
    # band edges for this radius, as relative distances
    t_in0, t_out0 = hj.filter_pair_band(kernel, d, eps, radius, rho, rho, 0.0, 0.0)
    w = max(np.sqrt(-2.0 * t_out0) / (s * eps) - 1.0,
            1.0 - np.sqrt(max(-2.0 * t_in0, 0.0)) / (s * eps))
    taus = np.concatenate([np.linspace(-3 * w, 3 * w, 25), np.linspace(-1.2 * w, 1.2 * w, 25)])
    decided_in = decided_out = 0
    for closer in (True, False):
        for ya, yb, delta, xa, xb in adversarial_frame_pairs(rng, d, s, eps, q, rho, taus, bf16,
                                                             closer):
            truth = oracle.pair_within(xa, xb, eps)
            ra, cb = slots(ya, fi, d, True, bf16), slots(yb, fi, d, False, bf16)
            dl = delta_slots(delta, fi, d)
            rho_a = round_f32(np.sqrt(np.sum(ya * ya)) * (1 + 2.0 ** -20))
            rho_b = round_f32(np.sqrt(np.sum(yb * yb)) * (1 + 2.0 ** -20))
            dd = [chunk_dd(dl[16 * t0:16 * t1], t1 - t0) for t0, t1 in chunks]
            dd_all = round_f32(sum(dd))
            t_in1, t_out1 = hj.filter_pair_band(kernel, d, eps, radius, rho_a, rho_b, dd[0], dd[0])
            t_in, t_out = hj.filter_pair_band(kernel, d, eps, radius, rho_a, rho_b, dd_all, dd_all)
            for dmode in ("fused", "seq", "pairfirst"):
                rc = [chunk_dot(ra[16 * t0:16 * t1], dl[16 * t0:16 * t1], t1 - t0, dmode)
                      for t0, t1 in chunks]
                cc = [chunk_dot(cb[16 * t0:16 * t1], dl[16 * t0:16 * t1], t1 - t0, dmode)
                      for t0, t1 in chunks]
                ra2 = corrected(ra, chunks, [-v for v in rc], parts, True, bf16)
                cb2 = corrected(cb, chunks, cc, parts, False, bf16)
                for order in ("seq", "rev", "tree", "once"):
                    acc = accumulate(ra2, cb2, order)
                    if acc >= t_in:
                        assert truth, (dmode, order, acc, t_in)
                        decided_in += 1
                    if acc < t_out:
                        assert not truth, (dmode, order, acc, t_out)
                        decided_out += 1
                    if steps > steps1:  # the early exit's partial test
                        acc1 = accumulate(ra2[:16], cb2[:16], order)
                        if acc1 < t_out1:
                            assert not truth
    assert decided_out > 0 and decided_in > 0


def test_frame_band_follows_the_tiles_radius():
    # the pair band scales with the tiles' own radii, not with the join's
    _, t_out_g = hj.filter_pair_band("f16", 28, 0.05, 2.65, 10854.0, 10854.0, 0.0, 0.0)
    _, t_out_t = hj.filter_pair_band("f16", 28, 0.05, 2.65, 400.0, 400.0, 0.0, 0.0)
    se = hj.filter_band("f16", 28, 0.05, 2.65)["scale"] * 0.05
    rel = lambda t: np.sqrt(-2.0 * t) / se - 1.0  # noqa: E731
    assert rel(t_out_t) < rel(t_out_g) / 10
    # the anchors' difference shifts the thresholds by |delta|^2 / 2
    t_in0, t_out0 = hj.filter_pair_band("f16", 28, 0.05, 2.65, 400.0, 400.0, 1e6, 0.0)
    t_in1, t_out1 = hj.filter_pair_band("f16", 28, 0.05, 2.65, 400.0, 400.0, 1e6, 1e6)
    assert abs((t_out1 - t_out0) - 5e5) < 5e5 * 1e-4
    with pytest.raises(hj.HJError):
        hj.filter_pair_band("fp64", 28, 0.05, 2.65, 1.0, 1.0, 0.0, 0.0)
