"""Seeded synthetic workloads for the direct-summation configurations.

Random numbers come from a counter-based SplitMix64 (Steele, Lea, Flood,
"Fast splittable pseudorandom number generators", OOPSLA 2014): element i of
stream `stream` under `seed` is mix64(key(seed, stream) + (i+1)*GOLDEN), so
any element can be produced independently, on the host or on the device.
Normals use Box-Muller on consecutive pairs.

Configurations (sizes as named):
  cube        N sources and N targets uniform in [0,1]^3, densities U[-1,1]
  sphere      sources uniform by area on the unit sphere with outward unit
              normals, targets uniform in the volume of the shell 1.05<r<2,
              densities N(0,1)
  clusters    points from 64 Gaussian clusters (sigma 0.02, centres uniform in
              [0,1]^3), 3-vector densities N(0,1)
  torus       near-field blocks on a torus (R=1, r=0.3) cut into nu x nv
              patches: per block a p x p Gauss-Legendre grid of sources with
              analytic normals (p = round(sqrt(Ns)), Ns log-uniform in
              [64,4096]) and Nt log-uniform in [16,2048] targets within 0.05 of
              the surface
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)
_SK = np.uint64(0xD1B54A32D192ED03)


def _mix64(z):
    z = (z ^ (z >> np.uint64(30))) * _M1
    z = (z ^ (z >> np.uint64(27))) * _M2
    return z ^ (z >> np.uint64(31))


def stream_key(seed, stream):
    with np.errstate(over="ignore"):
        return _mix64(np.array([np.uint64(seed) ^ (np.uint64(stream) * _SK)], dtype=np.uint64))[0]


def u64(seed, stream, n, offset=0):
    """n raw 64-bit draws of (seed, stream), elements offset..offset+n-1."""
    with np.errstate(over="ignore"):
        i = np.arange(offset + 1, offset + n + 1, dtype=np.uint64)
        return _mix64(stream_key(seed, stream) + i * GOLDEN)


def uniform(seed, stream, n, lo=0.0, hi=1.0):
    """n doubles uniform in [lo, hi): the top 53 bits of each draw."""
    x = (u64(seed, stream, n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return lo + (hi - lo) * x


def normal(seed, stream, n):
    """n N(0,1) doubles by Box-Muller over consecutive pairs of draws."""
    m = (n + 1) // 2
    x = u64(seed, stream, 2 * m)
    u1 = ((x[0::2] >> np.uint64(11)).astype(np.float64) + 1.0) * (1.0 / 9007199254740992.0)  # (0,1]
    u2 = (x[1::2] >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    rad = np.sqrt(-2.0 * np.log(u1))
    out = np.empty(2 * m)
    out[0::2] = rad * np.cos(2 * np.pi * u2)
    out[1::2] = rad * np.sin(2 * np.pi * u2)
    return out[:n]


# ------------------------------------------------------------------ configs
def cube(n_src, n_trg, seed=0, k0=1):
    xs = uniform(seed, 0, 3 * n_src)
    xt = uniform(seed, 1, 3 * n_trg)
    f = uniform(seed, 2, k0 * n_src, -1.0, 1.0)
    return xt, xs, f


def _unit_dirs(seed, stream, n):
    z = uniform(seed, stream, n, -1.0, 1.0)
    phi = uniform(seed, stream + 1, n, 0.0, 2 * np.pi)
    s = np.sqrt(np.maximum(0.0, 1.0 - z * z))
    return np.stack([s * np.cos(phi), s * np.sin(phi), z], axis=1)


def sphere(n_src, n_trg, seed=0, k0=1):
    """Returns (r_trg, r_src, n_src, v_src): sphere sources with n = x."""
    xs = _unit_dirs(seed, 0, n_src)
    d = _unit_dirs(seed, 2, n_trg)
    u = uniform(seed, 4, n_trg)
    r = np.cbrt(1.05 ** 3 + u * (2.0 ** 3 - 1.05 ** 3))
    xt = d * r[:, None]
    f = normal(seed, 5, k0 * n_src)
    return xt.ravel(), xs.ravel(), xs.ravel().copy(), f


def clusters(n_src, n_trg, seed=0, k0=3, n_clusters=64, sigma=0.02):
    centres = uniform(seed, 10, 3 * n_clusters).reshape(n_clusters, 3)

    def draw(n, s):
        c = np.minimum((uniform(seed, s, n) * n_clusters).astype(np.int64), n_clusters - 1)
        return (centres[c] + sigma * normal(seed, s + 1, 3 * n).reshape(n, 3)).ravel()

    xs = draw(n_src, 11)
    xt = draw(n_trg, 13)
    f = normal(seed, 15, k0 * n_src)
    return xt, xs, f


def torus_point(u, v, R=1.0, r=0.3):
    cv, sv, cu, su = np.cos(v), np.sin(v), np.cos(u), np.sin(u)
    x = np.stack([(R + r * cv) * cu, (R + r * cv) * su, r * sv], axis=-1)
    n = np.stack([cv * cu, cv * su, sv], axis=-1)
    return x, n


def torus_blocks(n_u=64, n_v=64, seed=0, k0=1, ns_range=(64, 4096), nt_range=(16, 2048), R=1.0, r=0.3,
                 offset=0.05):
    """Near-field blocks on a torus, one per (u, v) patch.

    Returns a dict with r_trg, r_src, n_src (flat AoS), v_src, and per-block
    trg_dsp/trg_cnt/src_dsp/src_cnt (blocks laid out consecutively)."""
    nb = n_u * n_v
    lns = np.log(ns_range)
    ns_draw = np.exp(lns[0] + (lns[1] - lns[0]) * uniform(seed, 20, nb))
    p = np.clip(np.rint(np.sqrt(ns_draw)).astype(np.int64), 8, 64)
    src_cnt = p * p
    lnt = np.log(nt_range)
    trg_cnt = np.rint(np.exp(lnt[0] + (lnt[1] - lnt[0]) * uniform(seed, 21, nb))).astype(np.int64)
    trg_cnt = np.clip(trg_cnt, nt_range[0], nt_range[1])
    src_dsp = np.concatenate([[0], np.cumsum(src_cnt)[:-1]])
    trg_dsp = np.concatenate([[0], np.cumsum(trg_cnt)[:-1]])
    du, dv = 2 * np.pi / n_u, 2 * np.pi / n_v
    gl = {}
    xs_all, ns_all, xt_all = [], [], []
    ut_all = uniform(seed, 22, int(trg_cnt.sum()))
    vt_all = uniform(seed, 23, int(trg_cnt.sum()))
    ht_all = uniform(seed, 24, int(trg_cnt.sum()), -offset, offset)
    for b in range(nb):
        iu, iv = b // n_v, b % n_v
        pb = int(p[b])
        if pb not in gl:
            gl[pb] = (np.polynomial.legendre.leggauss(pb)[0] + 1) / 2
        x = gl[pb]
        uu = (iu + x)[:, None] * du + 0 * x[None, :]
        vv = 0 * x[:, None] + (iv + x)[None, :] * dv
        X, N = torus_point(uu.ravel(), vv.ravel(), R, r)
        xs_all.append(X)
        ns_all.append(N)
        t0, tc = trg_dsp[b], trg_cnt[b]
        ut = (iu + ut_all[t0:t0 + tc]) * du
        vt = (iv + vt_all[t0:t0 + tc]) * dv
        Xt, Nt = torus_point(ut, vt, R, r)
        xt_all.append(Xt + ht_all[t0:t0 + tc, None] * Nt)
    r_src = np.concatenate(xs_all).ravel()
    n_src = np.concatenate(ns_all).ravel()
    r_trg = np.concatenate(xt_all).ravel()
    v_src = normal(seed, 25, k0 * int(src_cnt.sum()))
    return dict(r_trg=r_trg, r_src=r_src, n_src=n_src, v_src=v_src, trg_dsp=trg_dsp, trg_cnt=trg_cnt,
                src_dsp=src_dsp, src_cnt=src_cnt)


def near_field_case(n_u=8, n_v=6, n_trg=400, seed=0, k0=1, p_range=(4, 8), offset=0.05, stencil=1, R=1.0, r=0.3):
    """A BoundaryIntegralOp near field on a torus, as the reference's arrays.

    Elements are the n_u x n_v (u, v) patches, each with a p x p Gauss-Legendre
    tensor grid (p uniform in p_range) as its nodes; the far nodes are the
    nodes (the identity far-field density operator), with quadrature weights
    w_i w_j du dv |x_u x x_v| and analytic normals.  Targets are points of the
    torus moved along the normal by U[-offset, offset]; each target is near the
    (2*stencil+1)^2 patches around its own.  Returned keys follow
    boundary_integral.txx: X_far, Xn_far, wts_far, elem_nds_cnt, elem_nds_dsp
    (= *_far), Xtrg, Xn_trg, near_elem_cnt, near_elem_dsp, Xtrg_near,
    Xn_trg_near (element-major near pairs), near_scatter_index (target-major
    pair q <- element-major pair scatter[q]), near_trg_cnt, near_trg_dsp, and F
    (k0 N(0,1) densities per node)."""
    nelem = n_u * n_v
    p = p_range[0] + (u64(seed, 30, nelem) % np.uint64(p_range[1] - p_range[0] + 1)).astype(np.int64)
    du, dv = 2 * np.pi / n_u, 2 * np.pi / n_v
    X, Nn, W = [], [], []
    for e in range(nelem):
        iu, iv = e // n_v, e % n_v
        x, w = np.polynomial.legendre.leggauss(int(p[e]))
        x, w = (x + 1) / 2, w / 2
        uu = ((iu + x)[:, None] * du + 0 * x[None, :]).ravel()
        vv = (0 * x[:, None] + (iv + x)[None, :] * dv).ravel()
        Xe, Ne = torus_point(uu, vv, R, r)
        X.append(Xe)
        Nn.append(Ne)
        W.append((w[:, None] * w[None, :]).ravel() * du * dv * r * (R + r * np.cos(vv)))
    nds_cnt = p * p
    nds_dsp = np.concatenate([[0], np.cumsum(nds_cnt)[:-1]]).astype(np.int64)
    ut = uniform(seed, 31, n_trg) * 2 * np.pi
    vt = uniform(seed, 32, n_trg) * 2 * np.pi
    ht = uniform(seed, 33, n_trg, -offset, offset)
    Xt, Nt = torus_point(ut, vt, R, r)
    Xt = Xt + ht[:, None] * Nt
    iu_t = np.minimum((ut / du).astype(np.int64), n_u - 1)
    iv_t = np.minimum((vt / dv).astype(np.int64), n_v - 1)
    near = []
    for i in range(n_trg):
        es = {((iu_t[i] + a) % n_u) * n_v + (iv_t[i] + b) % n_v
              for a in range(-stencil, stencil + 1) for b in range(-stencil, stencil + 1)}
        near.append(sorted(es))
    pairs_e = [(e, i) for i in range(n_trg) for e in near[i]]
    order_em = sorted(range(len(pairs_e)), key=lambda q: (pairs_e[q][0], pairs_e[q][1]))
    em_index = np.empty(len(pairs_e), dtype=np.int64)
    em_index[np.array(order_em, dtype=np.int64)] = np.arange(len(pairs_e))
    near_trg = np.array([pairs_e[q][1] for q in order_em], dtype=np.int64)
    near_elem_cnt = np.bincount(np.array([pairs_e[q][0] for q in order_em], dtype=np.int64), minlength=nelem)
    near_elem_dsp = np.concatenate([[0], np.cumsum(near_elem_cnt)[:-1]]).astype(np.int64)
    near_trg_cnt = np.array([len(near[i]) for i in range(n_trg)], dtype=np.int64)
    near_trg_dsp = np.concatenate([[0], np.cumsum(near_trg_cnt)[:-1]]).astype(np.int64)
    Xf = np.concatenate(X)
    return dict(X_far=Xf.ravel(), Xn_far=np.concatenate(Nn).ravel(), wts_far=np.concatenate(W),
                elem_nds_cnt=nds_cnt.astype(np.int64), elem_nds_dsp=nds_dsp, elem_nds_cnt_far=nds_cnt.astype(np.int64),
                elem_nds_dsp_far=nds_dsp, Xtrg=Xt.ravel(), Xn_trg=Nt.ravel(),
                near_elem_cnt=near_elem_cnt.astype(np.int64), near_elem_dsp=near_elem_dsp,
                Xtrg_near=Xt[near_trg].ravel(), Xn_trg_near=Nt[near_trg].ravel(), near_scatter_index=em_index,
                near_trg_cnt=near_trg_cnt, near_trg_dsp=near_trg_dsp, F=normal(seed, 34, k0 * int(nds_cnt.sum())))


def interactions(blocks):
    return int(np.sum(blocks["trg_cnt"].astype(np.int64) * blocks["src_cnt"].astype(np.int64)))


# ----------------------------------------------------------------- metrics
def rel_errors(x, ref):
    """(normwise inf, RMS) relative errors: max|x-ref|/max|ref|, ||x-ref||_2/||ref||_2."""
    x = np.asarray(x, dtype=np.float64).ravel()
    ref = np.asarray(ref, dtype=np.float64).ravel()
    d = np.abs(x - ref)
    den_inf = np.abs(ref).max() if ref.size else 0.0
    den_2 = np.linalg.norm(ref)
    e_inf = float(d.max() / den_inf) if den_inf > 0 else float(d.max() if d.size else 0.0)
    e_2 = float(np.linalg.norm(x - ref) / den_2) if den_2 > 0 else float(np.linalg.norm(x - ref))
    return e_inf, e_2
