"""Throughput of the MI355X chain engine on its headline workload, one JSON line on stdout.

Headline: chain-levels/s of the batched complex-symmetric (route edc) recursion on the fcc ternary workload
(Fe0.66Ni0.14Cr0.20-like, spd, 2 shells = 18 slots, levels 4, rank 1, D2h-reduced node set, start site frozen to each
species, 2 spins: 9 x 3 x 2 = 54 chains of depth 30, FP64), next to the plain form (one chain per launch sequence, host
synchronisation per level) on the same chains, the KPM moments of the 40^3-cell fcc supercell (2,304,000 rows, 64
Rademacher vectors, 2048 moments), route ed3 on the same operator, the EPI pair blocks, the projected chains of the
spectral functions, the adjoint, and the oracle's CPU chains on the host cores.

A step is one batched pass of all 54 chains to depth 30; inputs (operator tables, start vectors) are resident on the GPU
when the timed region starts.  python bench.py [--steps K] [--warmup W] [--quick]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0       # MI355X HBM3E, GB/s
FP64_PEAK_TFS = 78.6        # MI355X dense FP64 (vector and matrix), TFLOP/s


def log(*a):
    print(*a, file=sys.stderr, flush=True)


# (vectors read, vectors written) per launch of a level of the route-edc chains (uniform d, o != 0, X(z) diagonal), by
# launch kind.  Separate update pass: 0 h1 = rand_c(x~) + D S D x~ (reads x; with the tile kernel it also writes the
# node-local t1 = X_o h1: then 6 writes t1 at the nodes with flip partners), [1 t1 = X_o h1], 0 w = sq (h1 - h t1 + X(z) x~)
# with the dot <u, w> (reads t1, h1, x), 3 the update (reads w, u, u_prev, writes w).  Fused level: 0 as above with the
# partials of a_n, 6 also the a_n terms of the listed nodes, 7 the second operator pass with the update (reads t1, h1, x and
# u_prev, writes the updated w).
EDC_IO_FIRST = (1, 2)
EDC_IO_SECOND = (3, 1)
EDC_IO_UPD = (4, 1)


def chain_kernel_cost(op, nb: int, kind: int, tile: int = 8, nread: int = 1, nwrite: int = 1):
    """Algorithmic bytes and flops of one launch of a chain-level kernel over nn nodes and nb complex members.

    Operator passes (kind 0, S pass + flip terms; kind 7, the same with the three-term update; kind 1, flip terms only):
    read nread vectors once (9 nb complex per node each), write nwrite once, read the node tables once (hop 4 nslot +
    flip 4 nsmax + info 4 + sq 8 bytes per node); flops: (1 + nslot) 9x9 blocks on complex members (4 flops per
    complex-by-real FMA pair) plus the flip terms and the D scalings (kind 7: the update, 9 complex a u + b u_prev and the
    norm).  Reductions (kinds 2, 4): the per-block partials, one complex value per block and member (plus the listed
    nodes' terms of a_n).  Update (3): w, u, u_prev read, w written.  Rescale (5): one vector read and written."""
    nn, ns, nsm = op.nn, op.nslot, op.nsmax
    vec = nn * 9 * nb * 16
    tables = nn * (4 * ns + 4 * nsm + 4 + 8)
    nbx = -(-nn // ((64 // min(tile, 64)) * 4))
    if kind in (0, 7):
        bytes_ = (nread + nwrite) * vec + tables
        flops = nn * nb * ((1 + ns) * 81 * 4 + nsm * 9 * 4 + (1 + ns) * 9 * 2 * 2 + 9 * 4 + (nwrite - 1) * 9 * 2)
        if kind == 7:
            flops += nn * nb * 9 * (8 + 8 + 4)
    elif kind == 1:
        bytes_ = (nread + 1) * vec + nn * (4 * nsm + 4 + 8)
        flops = nn * nb * (nsm * 9 * 4)
    elif kind == 3:
        bytes_ = 4 * vec
        flops = nn * nb * 9 * 16
    elif kind == 6:
        # the species-diagonal term at the nodes with flip partners only: own and partner rows read, one row written
        fr = np.asarray(op.flipr).reshape(nn, nsm)
        npf = int(np.sum(((fr >= 0) & (fr != np.arange(nn)[:, None])).any(axis=1)))
        bytes_ = npf * 9 * nb * 16 * (nsm + 1)
        flops = npf * nb * nsm * 9 * 4
    elif kind in (2, 4):
        bytes_ = nbx * -(-nb // tile) * tile * 16 + nb * 16 * 4
        flops = nbx * nb * 2
    else:
        bytes_ = 2 * vec
        flops = nn * nb * 9 * 6
    return float(bytes_), float(flops)


def level_io(kinds):
    """(vectors read, written) of each launch of a level from the sequence of launch kinds."""
    out, first = [], True
    for k in kinds:
        if k == 0:
            out.append(EDC_IO_FIRST if first and 6 in kinds else (EDC_IO_SECOND if not first else (1, 1)))
            first = False
        elif k == 7:
            out.append(EDC_IO_UPD)
        else:
            out.append((1, 1))
    return out


def compulsory_level_bytes(op, nb: int, s: int = 2):
    """The compulsory HBM bytes of one level of nb chains sharing one operator pass: per chain-level
    N [s 9 8 3 + T / n_share] (read u_k, read u_k-1, write u_k+1; the node tables T once per pass, shared by n_share = nb
    chains), T = the bytes per node of this engine's tables (hop 4 nslot + flip 4 nsmax + info 4 + sq 8)."""
    t = 4 * op.nslot + 4 * op.nsmax + 4 + 8
    return float(nb * op.nn * (s * 9 * 8 * 3) + op.nn * t)


def kpm_kernel_cost(nsite: int, nslot: int, nvec: int):
    """One Chebyshev step over nvec vectors: read t_n, read t_{n-1}, write t_{n+1} (the Rademacher vector is regenerated,
    never read); per site the species byte and nslot neighbour indices; flops: (1 + nslot) 9x9 blocks, the d scalings."""
    rows = nsite * 9
    bytes_ = rows * nvec * 8 * 3 + nsite * (1 + 4 * nslot)
    flops = nsite * nvec * ((1 + nslot) * 81 * 2 + (1 + nslot) * 9 + 9 * 6)
    return float(bytes_), float(flops)


def roofline(bytes_, flops, ms, mfma: bool = False):
    """Roofline of a launch: FP64-bound when its intensity exceeds the ridge (bound "fp64" for VALU FMA kernels, "mfma"
    for kernels that contract on the matrix cores; the FP64 peak of MI355X is 78.6 TFLOP/s for both), else HBM-bound."""
    t = ms * 1e-3
    gbs = bytes_ / t / 1e9
    tfs = flops / t / 1e12
    intensity = flops / bytes_
    if intensity > (FP64_PEAK_TFS * 1e12) / (HBM_PEAK_GBS * 1e9):
        return dict(bound="mfma" if mfma else "fp64", achieved=round(tfs, 3), peak=FP64_PEAK_TFS, unit="TFLOP/s",
                    frac=round(tfs / FP64_PEAK_TFS, 4), hbm_gbs=round(gbs, 1), hbm_frac=round(gbs / HBM_PEAK_GBS, 4),
                    flop_per_byte=round(intensity, 2))
    return dict(bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(gbs / HBM_PEAK_GBS, 4),
                fp64_tfs=round(tfs, 3), flop_per_byte=round(intensity, 2))


def measured_traffic(workload: str, kernel: str, nodes: int, members: int):
    """HBM bytes per launch of a kernel measured with rocprofv3 --pmc (FETCH_SIZE and WRITE_SIZE per dispatch, each scaled
    by the same counter's reading for a streaming kernel of known size: k_rescale, which reads and writes one chain vector),
    kept as text under profiles/ for the workload it was measured on; None when there is no measurement for this kernel
    and size."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as f:
            tab = json.load(f)
    except (OSError, ValueError):
        return None
    for rec in tab.get("kernels", []):
        if rec.get("workload") == workload and rec.get("kernel") == kernel and rec.get("nodes") == nodes and \
                rec.get("members") == members:
            return rec.get("hbm_bytes_per_launch")
    return None


def cpu_baseline(wl, nlev_sample: int, nchain: int | None = None):
    """The oracle's CPU chains (one chain per thread, dynamic schedule, as chain_worker) on the host cores: nchain chains
    (default: all of the workload) to nlev_sample levels each; cores = the oracle's OpenMP threads."""
    from oracle.oracle import OracleOp, threads
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    nch = len(wl.members) if nchain is None else min(len(wl.members), nchain)
    members = wl.members[:nch]
    eps = np.stack([m.z * m.pa + m.pb for m in members])
    O = OracleOp(wl.op)
    t0 = time.perf_counter()
    a, b, mu = O.edc_batch(members, eps, [m.l0 for m in members], [nlev_sample] * nch, project=True)
    dt = time.perf_counter() - t0
    return dict(value=round(float(np.sum(mu)) / dt, 3), unit="chain-levels/s", cores=threads(), kind="port",
                sample=f"{nch} chains x {nlev_sample} levels of the same workload ({wl.op.nn} nodes), {dt:.1f} s"), (a, b, mu)


def kpm_cpu_baseline(g, species, par, block, block0, cshift, sscale, nvec: int, nmom: int, per_gpu):
    """The oracle's KPM moments (one vector after another, each step a parallel row loop over the sites as csr_apply) on
    the host cores: the first nvec Rademacher vectors of the same seed x nmom moments of the same supercell.  per_gpu
    (nvec x nmom) holds the GPU's per-vector terms <r|t_n>/dim of the same vectors: the largest difference is rounding."""
    from oracle.oracle import OracleReal, threads
    os.environ.setdefault("OMP_PROC_BIND", "spread")
    O = OracleReal(g.nbr, species, par[0], par[1], par[2], par[3], block, block0)
    t0 = time.perf_counter()
    _, per = O.kpm_moments(cshift, sscale, nmom, nvec, 5, per_vector=True)
    dt = time.perf_counter() - t0
    return dict(value=round(nvec * nmom / dt, 3), unit="vector-moments/s", cores=threads(), kind="port",
                sample=f"{nvec} vectors x {nmom} moments of the same supercell ({g.n * 9} rows), {dt:.1f} s",
                max_abs_diff_per_vector_vs_gpu=float(np.max(np.abs(per - per_gpu))))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--quick", action="store_true", help="small sizes (smoke)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-kpm", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--kpm-moments", type=int, default=2048)
    ap.add_argument("--no-l21", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true")
    ap.add_argument("--no-ed3", action="store_true")
    ap.add_argument("--no-d4h", action="store_true")
    ap.add_argument("--no-epi", action="store_true")
    ap.add_argument("--no-spectral", action="store_true")
    ap.add_argument("--no-scf", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: the engine's multi-GPU form (chains dealt across GPUs, one all-gather of the "
              "coefficients) is not part of this build; run with --gpus 1", file=sys.stderr)
        sys.exit(2)

    from asr_amd import engine, workloads as W
    if engine.device_count() < 1:
        raise SystemExit("no GPU visible")
    L, levels, depth = (4, 2, 30) if args.quick else (8, 4, 30)
    t = time.perf_counter()
    wl = W.ternary_edc(L=L, levels=levels, depth=depth)
    log(f"workload: {wl.name}, {wl.op.nn} nodes, {len(wl.members)} chains x {depth}, build {time.perf_counter() - t:.1f} s")
    G = engine.GpuOperator(wl.op)
    nb = len(wl.members)
    chain_levels = int(np.sum(wl.nlev))

    for _ in range(args.warmup):
        G.chains_edc(wl.members, wl.nlev)
    dev_ms = []
    t0 = time.perf_counter()
    for _ in range(args.steps):
        a, b, mu, ms = G.chains_edc(wl.members, wl.nlev)
        dev_ms.append(ms)
    wall = time.perf_counter() - t0
    ms_per_step = wall / args.steps * 1e3
    value = chain_levels / (wall / args.steps)
    log(f"batched: {ms_per_step:.1f} ms/step (device {np.mean(dev_ms):.1f} ms), {value:.0f} chain-levels/s")

    out_tile = G.tuned_tile(nb, True)
    log(f"autotuned member tile: {out_tile}")
    # per-kernel durations of the level step (HIP events on the engine's stream)
    kms, kinds = G.profile(wl.members, wl.nlev)
    kinds = [int(k) for k in kinds]
    io = level_io(kinds)
    per_kernel = []
    for q, (t_ms, kd) in enumerate(zip(kms, kinds)):
        by, fl = chain_kernel_cost(wl.op, nb, kd, out_tile, *io[q])
        per_kernel.append(dict(launch=q, kind=kd, ms=round(float(t_ms), 4), gbs=round(by / (t_ms * 1e-3) / 1e9, 1),
                               tfs=round(fl / (t_ms * 1e-3) / 1e12, 3)))
    # the dominant kernel is the operator pass.  With the tile kernel its two launches per level are two instantiations
    # (k_tile<..., 1>: the first S pass, which also writes the node-local o term and, in the fused level, the partials of
    # a_n; k_tile<..., 0>: the second S pass with the output terms and, in the fused level, the three-term update), so each
    # gets its own roofline, the longer one as "roofline"; otherwise every launch of kinds 0 and 1 is the same kernel and
    # the roofline is over their summed cost and duration
    piid = bool(np.all(np.asarray(wl.op.pi) == np.arange(9)[None, :]))
    tile_kernel = out_tile in (8, 16) and piid and wl.op.nslot in (14, 18)
    split = 6 in kinds

    def launch_roofline(launches, kname):
        tb = tf = tm = 0.0
        for k in launches:
            by, fl = chain_kernel_cost(wl.op, nb, k["kind"], out_tile, *io[k["launch"]])
            tb, tf, tm = tb + by, tf + fl, tm + k["ms"]
        r = roofline(tb / len(launches), tf / len(launches), tm / len(launches))
        r["kernel"] = kname
        r["launches_per_level"] = len(launches)
        r["avg_launch_ms"] = round(tm / len(launches), 4)
        r["algorithmic_bytes_per_launch"] = round(tb / len(launches))
        r["traffic"] = measured_traffic("chain_edc", kname, wl.op.nn, nb)
        return r

    roof_first = None
    if tile_kernel and split:
        ops = [k for k in per_kernel if k["kind"] in (0, 7)]
        first, second = ops[0], ops[-1]
        roof = launch_roofline([second], "k_tile<cplx, %d, %d, 0>" % (out_tile, wl.op.nslot))
        roof_first = launch_roofline([first], "k_tile<cplx, %d, %d, 1>" % (out_tile, wl.op.nslot))
    else:
        ops = [k for k in per_kernel if k["kind"] in (0, 1)]
        if tile_kernel:
            kname = "k_tile<cplx, %d, %d, 0>" % (out_tile, wl.op.nslot)
        else:
            kname = {65: "k_pair<true, %d>" % wl.op.nslot, 64: "k_node<cplx, true, %d>" % wl.op.nslot}.get(out_tile,
                                                                                                         "k_fused<cplx>")
        roof = launch_roofline(ops, kname)
    # the whole level against the compulsory traffic of the recursion (read u_k, u_k-1, write u_k+1, the tables once)
    lev_ms = float(np.mean(dev_ms)) / depth
    lev_b = compulsory_level_bytes(wl.op, nb)
    roof_level = dict(bound="hbm", achieved=round(lev_b / (lev_ms * 1e-3) / 1e9, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                      frac=round(lev_b / (lev_ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), level_ms=round(lev_ms, 4),
                      compulsory_bytes_per_level=round(lev_b), launches_per_level=len(kinds),
                      form="fused level (a_n from the first pass, update in the second)" if 7 in kinds else
                      "separate update pass")
    log("level kernels:", per_kernel)

    out = {
        "metric": "chain-levels/s (batched complex-symmetric ASR chains, 1 GPU)",
        "value": round(value, 1),
        "unit": "chain-levels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 2),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded species parameters, screened fcc structure constants)",
        "config": {"workload": "fcc ternary 0.66/0.14/0.20, route edc, levels 4, rank 1, D2h, 54 chains x 30",
                   "nodes": wl.op.nn, "chains": nb, "depth": depth, "L": L, "levels": levels,
                   "device_ms_per_step": round(float(np.mean(dev_ms)), 2), "member_tile": out_tile,
                   "node_order": wl.op.meta.get("node_order", "bfs")},
        "roofline": roof,
        "roofline_first_pass": roof_first,
        "roofline_level": roof_level,
        "level_kernels": per_kernel,
    }

    if not args.no_plain:
        t0 = time.perf_counter()
        a1, b1, mu1, ms1 = G.chains_edc(wl.members, wl.nlev, mode=1)
        wall1 = time.perf_counter() - t0
        plain = chain_levels / wall1
        out["plain"] = {"value": round(plain, 1), "unit": "chain-levels/s", "ms": round(wall1 * 1e3, 1),
                        "form": "one chain per launch sequence, host sync per level"}
        out["speedup_batched_vs_plain"] = round(value / plain, 2)
        out["plain_matches_batched"] = bool(np.max(np.abs(a1 - a)) <= 1e-12 * np.max(np.abs(a)))
        log(f"plain: {wall1 * 1e3:.0f} ms, {plain:.0f} chain-levels/s, batched/plain {value / plain:.2f}x")

    if not args.no_d4h:
        # the reference's own point group (class_ops' D4h, which swaps p_x / p_y and d_xz / d_yz): the same fcc ternary
        # reduced by D4h (about half the nodes of D2h), the 30 chains whose start orbital is one-dimensional under the
        # group (2 spins x 3 central species x s, p_z, d_z2, d_xy, d_x2-y2), depth 30; batched and plain
        wd = W.ternary_edc(L=L, levels=levels, depth=depth, group="d4h")
        md4 = [m for m in wd.members if m.l0 in (0, 2, 4, 6, 8)]
        G4 = engine.GpuOperator(wd.op)
        G4.chains_edc(md4, depth)  # warm-up (autotuning)
        res = {}
        for mode, name in ((0, "batched"), (1, "plain")):
            t0 = time.perf_counter()
            a4, b4, mu4, ms4 = G4.chains_edc(md4, depth, mode=mode)
            wall = time.perf_counter() - t0
            res[name] = {"value": round(len(md4) * depth / wall, 1), "wall_ms": round(wall * 1e3, 1),
                         "device_ms": round(ms4, 1)}
        out["d4h"] = {"workload": "fcc ternary, route edc, levels 4, D4h (orbital-permuting), 30 chains x 30",
                      "nodes": wd.op.nn, "chains": len(md4), "depth": depth, "unit": "chain-levels/s", **res,
                      "member_tile": G4.tuned_tile(len(md4), True),
                      "speedup_batched_vs_plain": round(res["batched"]["value"] / res["plain"]["value"], 2)}
        log(f"d4h: {out['d4h']}")
        G4.close()

    if not args.no_ed3:
        # route ed3 on the same operator: 54 (spin, central species, start orbital) blocks, two chains each, depth 30
        we = W.ternary_ed3(L=L, levels=levels, depth=depth, op=wl.op)
        cf, oc = we.info["cf"], we.info["oc"]
        G.set_depth_tol(0.0, "hull")
        G.ed3_blocks(we.members, cf, oc, we.nlev)  # warm-up (autotuning)
        res = {}
        for mode, name in ((0, "batched"), (2, "graph"), (1, "plain")):
            t0 = time.perf_counter()
            bk, a3, b3, c3, mu3, ms3 = G.ed3_blocks(we.members, cf, oc, we.nlev, mode=mode)
            wall = time.perf_counter() - t0
            lev = int(mu3.sum())
            res[name] = {"value": round(lev / wall, 1), "wall_ms": round(wall * 1e3, 1), "device_ms": round(ms3, 1)}
        out["ed3"] = {"workload": "fcc ternary, route ed3 (full-ASA energy-dependent), 54 blocks x 2 chains x 30, "
                                  "the edc operator", "nodes": wl.op.nn, "chains": 2 * len(we.members), "depth": depth,
                      "unit": "chain-levels/s", **res, "member_tile": G.tuned_tile(2 * len(we.members), True),
                      "speedup_batched_vs_plain": round(res["batched"]["value"] / res["plain"]["value"], 2)}
        log(f"ed3: {out['ed3']}")

    if not args.no_epi:
        # EPI pair blocks: first-shell pair on the fcc ternary (levels 3, D2h stabiliser of the partner), nine adapted
        # start vectors x two partner species, two chains each, depth 30
        t = time.perf_counter()
        wp = W.fcc_epi_pair(L=4, levels=2) if args.quick else W.fcc_epi_pair(L=8, levels=3)
        GP = engine.GpuOperator(wp.op)
        log(f"epi: {wp.op.nn} nodes, {len(wp.members)} blocks, build {time.perf_counter() - t:.1f} s")
        GP.epi_blocks(wp.members, wp.info["v9"], 30)  # warm-up (autotuning)
        res = {}
        for mode, name in ((0, "batched"), (2, "graph"), (1, "plain")):
            t0 = time.perf_counter()
            x, _, _, _, mu3, ms3 = GP.epi_blocks(wp.members, wp.info["v9"], 30, mode=mode)
            wall = time.perf_counter() - t0
            res[name] = {"value": round(int(mu3.sum()) / wall, 1), "wall_ms": round(wall * 1e3, 1), "device_ms": round(ms3, 1)}
        out["epi"] = {"workload": "fcc ternary first-shell EPI pair, levels 3, 18 blocks x 2 chains x 30", "nodes": wp.op.nn,
                      "chains": 2 * len(wp.members), "unit": "chain-levels/s", **res,
                      "member_tile": GP.tuned_tile(2 * len(wp.members), True),
                      "speedup_batched_vs_plain": round(res["batched"]["value"] / res["plain"]["value"], 2)}
        log(f"epi: {out['epi']}")
        GP.close()

    if not args.no_spectral:
        # projected chains (spectral functions): the headline chains with their projections on the 19 sites of the
        # start site's first two shells at every level
        t = time.perf_counter()
        ws = W.ternary_edc(L=L, levels=levels, depth=depth, targets=True)
        GS = engine.GpuOperator(ws.op)
        GS.set_targets(ws.op.meta["tnode"], ws.op.meta["tgop"])
        log(f"spectral: {ws.op.nn} nodes, {len(ws.members)} chains x {len(ws.info['tsite'])} targets, build "
            f"{time.perf_counter() - t:.1f} s")
        GS.chains_edc_proj(ws.members, ws.nlev, ws.info["tw"])  # warm-up (autotuning)
        res = {}
        for mode, name in ((0, "batched"), (2, "graph"), (1, "plain")):
            t0 = time.perf_counter()
            r = GS.chains_edc_proj(ws.members, ws.nlev, ws.info["tw"], mode=mode)
            wall = time.perf_counter() - t0
            res[name] = {"value": round(int(r[5].sum()) / wall, 1), "wall_ms": round(wall * 1e3, 1),
                         "device_ms": round(r[6], 1)}
        out["spectral"] = {"workload": "fcc ternary projected edc chains, 54 chains x 30, 19 target sites",
                           "nodes": ws.op.nn, "chains": len(ws.members), "targets": len(ws.info["tsite"]),
                           "unit": "chain-levels/s", **res,
                           "speedup_batched_vs_plain": round(res["batched"]["value"] / res["plain"]["value"], 2)}
        log(f"spectral: {out['spectral']}")
        GS.close()

    if not args.no_scf:
        # one occupation pass of the SCF loop on the headline operator (chain_greens + ed_moments): the Fermi contour
        # with the reference's defaults (32 main-line nodes, 16 across the step, kT = 1 mRy: 16 poles) plus the E_F probe,
        # every (energy, start orbital, component, spin) task a chain of depth 30 with the adaptive depth (1e-5, hull),
        # in energy-major batches of at most 112 members that reuse the operator's cached level vectors
        from asr_amd import scf as SC
        params = wl.info["params"]
        enu = np.stack([np.stack([p["enu"][pp] for p in params]) for pp in range(3)])
        cp = SC.contour(0.0, enu)
        classes, comps = [SC.ChainClass(wl.op)], [SC.Component(0) for _ in range(3)]
        # the same pass at fixed depth (every chain to 30 levels, in batches of 112), for the saving of the adaptive depth
        t0 = time.perf_counter()
        cg0 = SC.chain_greens(cp.zq, classes, comps, params, depth, route="edc", depth_tol=0.0, gpu_ops=[G])
        wall0 = time.perf_counter() - t0
        t0 = time.perf_counter()
        cg = SC.chain_greens(cp.zq, classes, comps, params, depth, route="edc", depth_tol=1.0e-5, tail="hull",
                             gpu_ops=[G])
        mom = SC.ed_moments(cp, cg, classes, comps, enu)
        wall = time.perf_counter() - t0
        G.set_depth_tol(0.0, "hull")
        out["scf_pass"] = {"workload": "fcc ternary edc occupation pass: Fermi contour (kT 1 mRy, 32 + 16 + 4 + 8 nodes, "
                                       "16 poles) + E_F probe, 2 spins x 3 components x 9 orbitals, depth 30, "
                                       "adaptive depth 1e-5", "nodes": wl.op.nn, "energies": int(cp.zq.size),
                           "chains": cg.nchain, "chain_levels": cg.nlev, "unit": "chain-levels/s",
                           "value": round(cg.nlev / wall, 1), "wall_s": round(wall, 2),
                           "device_s": round(cg.device_ms * 1e-3, 2),
                           "form": "one window run of 112 member slots (queued tasks take the slots of chains that "
                                   "leave), the depth test on its own stream",
                           "fixed_depth": {"chain_levels": cg0.nlev, "wall_s": round(wall0, 2),
                                           "device_s": round(cg0.device_ms * 1e-3, 2)},
                           "chain_levels_saved": round(1.0 - cg.nlev / cg0.nlev, 3),
                           "device_time_saved": round(1.0 - cg.device_ms / cg0.device_ms, 3),
                           "wall_time_saved": round(1.0 - wall / wall0, 3),
                           "causality_failures": [mom.npsd, mom.nnode],
                           "charge_per_spin": [round(float(mom.qc[:, g].sum()), 6) for g in range(2)]}
        log(f"scf pass: {out['scf_pass']}")

    if not args.no_l21:
        # uneven chains over four sublattice operators (L2_1 partial disorder + SRO) in one batched run
        t = time.perf_counter()
        wls = W.l21_batch(L=4, levels=(2, 2, 2, 1), nchains=48, depth=(5, 12)) if args.quick else W.l21_batch()
        groups = [(engine.GpuOperator(w.op), w.members, w.nlev) for w in wls]
        tot_lev = int(sum(int(np.sum(w.nlev)) for w in wls))
        log(f"l21: {[(w.op.nn, len(w.members)) for w in wls]} nodes/chains, {tot_lev} chain-levels, build "
            f"{time.perf_counter() - t:.1f} s")
        engine.chains_multi(groups, mode=2)  # warm-up (autotuning)
        res = {}
        for mode, name in ((2, "graph"), (0, "streams"), (1, "plain")):
            t0 = time.perf_counter()
            _, dms = engine.chains_multi(groups, mode=mode)
            wall = time.perf_counter() - t0
            res[name] = {"value": round(tot_lev / wall, 1), "wall_ms": round(wall * 1e3, 1), "device_ms": round(dms, 1)}
        out["batched_uneven"] = {"workload": "L2_1 4 sublattices (1.0 / 0.5-0.5 / 0.4-0.6 / 0.3-0.3-0.4), SRO alpha_1, "
                                             "2 spins, 256 chains of depth 20-60, levels (3, 3, 3, 2)",
                                 "nodes": [w.op.nn for w in wls], "chains": [len(w.members) for w in wls],
                                 "chain_levels": tot_lev, "unit": "chain-levels/s", **res,
                                 "speedup_graph_vs_plain": round(res["graph"]["value"] / res["plain"]["value"], 2)}
        log(f"l21: {res}")
        for G, _, _ in groups:
            G.close()

    if not args.no_kpm:
        Lk = 6 if args.quick else 40
        nmom = 64 if args.quick else args.kpm_moments
        g, species, par, block, block0 = W.kpm_supercell(L=Lk)
        R = engine.GpuRealization(g.nbr, species, par, block, block0)
        # spectral bounds of h = c + d S d from 80 Lanczos steps (spectral_bounds), widened by 1 % of the width on each
        # side, and the spectrum scaled into [-0.99, 0.99]
        t0 = time.perf_counter()
        elo, ehi, _, _ = R.spectral_bounds(80, margin=0.0, nvec=1, seed=3)
        tb_ms = (time.perf_counter() - t0) * 1e3
        marg = 0.01 * (ehi - elo)
        cshift, sscale = 0.5 * (elo + ehi), 0.5 * (ehi - elo + 2.0 * marg) / 0.99
        tiles = {}
        # one or two vectors per lane on 64-vector rows, 16-vector slices, or the block contraction on the FP64 matrix
        # cores; tuned on short runs
        for vt in (1, 2, 3, 4):
            R.set_tile(vt)
            tiles[vt] = R.kpm_moments(cshift, sscale, 12, 64, seed=5)["step_ms"]
        vt_best = min(tiles, key=tiles.get)
        R.set_tile(vt_best)
        log(f"kpm vector tiles (ms/step): {tiles} -> {vt_best}")
        t0 = time.perf_counter()
        res = R.kpm_moments(cshift, sscale, nmom, 64, seed=5)
        wk = time.perf_counter() - t0
        # the Jackson-kernel DOS on 4001 points inside the scaled window (kpm_dos, on the host)
        from asr_amd import host as H
        t0 = time.perf_counter()
        egrid = cshift + sscale * 0.99 * np.linspace(-1.0, 1.0, 4001)
        rho = H.kpm_dos(res["mu"], cshift, sscale, egrid)
        tdos = time.perf_counter() - t0
        dos_norm = float(np.sum(rho) * (egrid[1] - egrid[0]))
        kby, kfl = kpm_kernel_cost(g.n, g.nbr.shape[1], 64)
        kroof = roofline(kby, kfl, res["step_ms"])
        kroof["kernel"] = {1: "k_kpm_u<false, %d>", 2: "k_kpm_v2<%d, false>", 3: "k_kpm_s16<%d>"}[vt_best] % g.nbr.shape[1]
        kroof["algorithmic_bytes_per_launch"] = round(kby)
        kroof["traffic"] = measured_traffic("kpm", kroof["kernel"], g.n, 64)
        out["kpm"] = {"value": round(64 * nmom / wk, 1), "unit": "vector-moments/s", "rows": g.n * 9, "nvec": 64,
                      "nmom": nmom, "wall_s": round(wk, 3), "step_ms": round(res["step_ms"], 4), "roofline": kroof,
                      "vector_tile": vt_best,
                      "bounds": {"emin": round(elo - marg, 6), "emax": round(ehi + marg, 6), "lanczos_steps": 80,
                                 "ms": round(tb_ms, 1)},
                      "mu_head": [round(float(v), 8) for v in res["mu"][:4]],
                      "dos": {"points": 4001, "ms": round(tdos * 1e3, 2), "integral": round(dos_norm, 6)}}
        log(f"kpm: {nmom} moments x 64 vectors on {g.n * 9} rows in {wk:.2f} s, step {res['step_ms']:.3f} ms, {kroof}")
        # the same moments by moment doubling (mu_2m = 2 <t_m|t_m> - mu_0, mu_2m-1 = 2 <t_m|t_m-1> - mu_1): nmom / 2
        # Chebyshev steps, no random vector in the step; the difference from the direct moments is rounding
        R.set_doubling(True)
        R.kpm_moments(cshift, sscale, 12, 64, seed=5)
        t0 = time.perf_counter()
        rd = R.kpm_moments(cshift, sscale, nmom, 64, seed=5)
        wd = time.perf_counter() - t0
        R.set_doubling(False)
        dby = g.n * 9 * 64 * 8 * 3.0  # read t_m (its 19 gathers counted once), read t_m-1, write t_m+1
        droof = roofline(dby, kfl, rd["step_ms"])
        droof["kernel"] = "k_kpm_v2<%d, true>" % g.nbr.shape[1]
        droof["algorithmic_bytes_per_launch"] = round(dby)
        droof["traffic"] = measured_traffic("kpm", droof["kernel"], g.n, 64)
        out["kpm"]["doubling"] = {"value": round(64 * nmom / wd, 1), "unit": "vector-moments/s", "wall_s": round(wd, 3),
                                  "steps": nmom // 2 - 1, "step_ms": round(rd["step_ms"], 4), "roofline": droof,
                                  "max_abs_diff_vs_direct": float(np.max(np.abs(rd["mu"] - res["mu"])))}
        log(f"kpm doubling: {out['kpm']['doubling']}")
        if not args.no_cpu:
            nvc, nmc = (2, 16) if args.quick else (4, 64)
            per = R.kpm_moments(cshift, sscale, nmc, 64, seed=5, per_vector=True)["mu_per"][:nvc]
            out["kpm"]["cpu_baseline"] = kpm_cpu_baseline(g, species, par, block, block0, cshift, sscale, nvc, nmc, per)
            log(f"kpm cpu baseline: {out['kpm']['cpu_baseline']}")
        R.close()

    if not args.no_adjoint:
        # gradients of a contour objective through 256 real chains of depth 40 (18 parameters, checkpoints every 8)
        from asr_amd import adjoint as AD
        import torch
        t = time.perf_counter()
        wa = W.ternary_adjoint(L=4, levels=2, nchains=32, depth=16) if args.quick else W.ternary_adjoint()
        nla, nca = int(wa.nlev[0]), len(wa.members)
        log(f"adjoint: {wa.op.nn} nodes, {nca} chains x {nla}, build {time.perf_counter() - t:.1f} s")
        GA = engine.GpuOperator(wa.op)
        mem = AD.members_at(wa.members, wa.info["theta"], wa.info["c_offset"])
        rng = np.random.default_rng(9)
        la, lb = rng.standard_normal((nca, nla)), rng.standard_normal((nca, nla))
        GA.lanczos(mem, nla, tiny=0.0)  # warm-up (autotuning)
        _, _, _, fwd_plain = GA.lanczos(mem, nla, tiny=0.0)
        GA.lanczos_adjoint(mem, nla, la, lb)
        t0 = time.perf_counter()
        _, _, gr, (fms, ams) = GA.lanczos_adjoint(mem, nla, la, lb)
        wadj = time.perf_counter() - t0
        # the torch harness end to end: band energy + charge penalty on a 64-node closed contour
        from asr_amd import host as H
        z, w = H.contour_closed(-1.6, -0.05, 0.6, 48, 16)
        weight = torch.full((nca,), 1.0 / nca, dtype=torch.float64)
        q0 = torch.full((nca,), 0.5, dtype=torch.float64)
        theta = torch.tensor(wa.info["theta"], requires_grad=True)
        t0 = time.perf_counter()
        a_t, b_t = AD.chain_coefficients(theta, GA, wa.members, nla, wa.info["c_offset"])
        lossv = AD.contour_objective(a_t, b_t, torch.from_numpy(z), torch.from_numpy(w), weight, q0)
        t1 = time.perf_counter()
        lossv.backward()
        t2 = time.perf_counter()
        out["adjoint"] = {"workload": "fcc ternary, real chains on h = c + d S d, 18 parameters (c, d x 3 l x 3 species), "
                                      "checkpoints every 8 levels", "nodes": wa.op.nn, "chains": nca, "depth": nla,
                          "forward_ms": round(fwd_plain, 2), "forward_ckpt_ms": round(fms, 2), "reverse_ms": round(ams, 2),
                          "reverse_over_forward": round(ams / fwd_plain, 2),
                          # the whole gradient (the checkpointed forward that the reverse sweep needs, then the reverse)
                          # against one plain forward: the figure a <= 3x budget for a gradient is judged by
                          "gradient_over_forward": round((fms + ams) / fwd_plain, 2), "wall_ms": round(wadj * 1e3, 1),
                          "torch_forward_s": round(t1 - t0, 3), "torch_backward_s": round(t2 - t1, 3),
                          "grad_norm": float(np.linalg.norm(gr.sum(axis=0)))}
        log(f"adjoint: {out['adjoint']}")
        GA.close()

    if not args.no_cpu:
        cb, _ = cpu_baseline(wl, 2 if args.quick else depth)
        out["cpu_baseline"] = cb
        log(f"cpu baseline: {cb}")
    G.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
