"""Benchmark of the CBM product on one MI355X.

A step is one forward product Y = M X plus one adjoint dX = M^T G over the
full synthetic graph of the workload, inputs resident in HBM, with M the
workload's product: A (cbm4mm: c1, c5), A D^-1/2 (cbm4ad: c4) or
D^-1/2 A D^-1/2 (cbm4dad: c2, c3).  Reported on the JSON line:

  value            ms per step, BF16 storage of X and G, FP32 accumulation
  detail           the FP32 step, the two halves of each product, the
                   rocSPARSE FP32 CSR step on the same GPU (fastest of its
                   SpMM algorithms, A^T as an explicit CSR), the same
                   gathers without compression (our kernel over plain A)
  roofline         the dominant phase of the step against 8 TB/s HBM
  cpu_baseline     torch.sparse CSR (MKL) FP32 M@X + M^T@G on the host cores
  secondary        the compressible workload (c4) with the same fields
  build_stages     the structure build by stage: adjacency CSR, candidate
                   scan on the GPU (HIP kernels) vs on the host cores,
                   arborescence + Delta

    python bench.py [--steps K] [--warmup W] [--workload c5] [--alpha 8] [--cols 128]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cbm-gpu_amd"))

HBM_PEAK_GBS = 8000.0

# Seeded community copying model per workload: N, blocks, copy probability,
# random-link rate lambda, first-node degree of a block, graph seed, alpha, F,
# and the product.  lambda (and the seed degree) are set so that nnz lands on
# the stated size.
WORKLOADS = {
    "c1": dict(n=2708, blocks=7, p=0.8, lam=0.58, seed_deg=0, seed=1001, alpha=0, cols=64, product="mm"),
    "c2": dict(n=19717, blocks=3, p=0.85, lam=0.44, seed_deg=0, seed=1002, alpha=8, cols=128, product="dad"),
    "c3": dict(n=169343, blocks=40, p=0.8, lam=1.63, seed_deg=0, seed=1003, alpha=16, cols=128, product="dad"),
    "c4": dict(n=132534, blocks=8, p=0.95, lam=20.0, seed_deg=225, seed=1004, alpha=32, cols=256, product="ad"),
    "c5": dict(n=2449029, blocks=47, p=0.75, lam=6.6, seed_deg=0, seed=1005, alpha=8, cols=128, product="mm"),
}

# rocsparse_spmm_alg values tried for the CSR comparator.
ROCSPARSE_ALGS = {"default": 0, "csr_row_split": 4, "csr_nnz_split": 5, "csr_merge_path": 9}


def measured_traffic(workload, phase):
    """Memory bytes per call of a phase from the committed counter summary
    (profiles/hbm_traffic.json, written from rocprofv3 runs of the 32-byte
    DRAM read-request counter and WRITE_SIZE, each checked against a copy of
    known size): bytes that leave the L2s, or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
        return data[workload][phase]
    except (OSError, KeyError, ValueError):
        return None


def committed_gather_ceiling(workload, rec):
    """The rate of the same Delta-pass kernel over uniform-random indices of
    the workload's shape (bench.py --gather-ceiling, committed in
    profiles/gather_ceiling.json), in gathered GB/s, or None."""
    path = os.path.join(ROOT, "profiles", "gather_ceiling.json")
    try:
        with open(path) as fh:
            data = json.load(fh)
        c = data[workload]
        return {"gbs": c["above_mall"]["gather_gbs"], "source": f"committed profiles/gather_ceiling.json ({c['sample']})"}
    except (OSError, KeyError, ValueError):
        return None


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def phase_bytes(n, nnz_delta, n_edges, n_internal, depth, f, s_in, scaled, symmetric):
    """Algorithmic (compulsory) HBM bytes of each phase: every structure
    array read once, every input row read once, every output row written
    once; gathered rows read again, and intermediates, are not counted.
    `scaled`: the product carries the D^-1/2 vector (4N bytes).  With a
    symmetric A the adjoint runs through the forward structure, so its bytes
    are the forward's (the Delta^T form's are given for its two halves).
    `gather_*`: the bytes of the row gathers themselves (nnz(Delta) rows of
    s_in * F bytes), which no cache on the chip holds at the largest sizes."""
    dvec = 4 * n if scaled else 0
    fwd = 4 * (n + 1) + 4 * nnz_delta + 8 * n_edges + 4 * (depth + 1) + s_in * n * f + 4 * n * f + dvec
    delta = 4 * (n + 1) + 4 * nnz_delta + s_in * n * f + 4 * n * f + dvec
    rev = 4 * (n_internal + 1) + 4 * n_internal + 4 * n_edges + s_in * f * (n_internal + n_edges) + 4 * f * n_internal
    dlt = 4 * (n + 1) + 4 * n + 4 * nnz_delta + s_in * n * f + 4 * n_internal * f + 4 * n * f + dvec
    step_bwd = 8 * (n + 1) + 4 * nnz_delta + 4 * n_edges + 4 * (depth + 1) + s_in * n * f + 4 * n * f + dvec
    return {"forward": fwd, "forward_delta_spmm": delta, "adjoint_reverse": rev, "adjoint_delta_t": dlt,
            "step": fwd + (fwd if symmetric else step_bwd), "gather_delta": nnz_delta * s_in * f}


def csr_bytes(n, nnz, f):
    return 4 * (n + 1) + 8 * nnz + 8 * n * f


CBM_CLASSES = {"mm": ("cbm.cbm4mm", "cbm4mm"), "ad": ("cbm.cbm4ad", "cbm4ad"), "dad": ("cbm.cbm4dad", "cbm4dad")}
CSR_CLASSES = {"mm": ("cbm.csr4mm", "csr4mm"), "ad": ("cbm.csr4ad", "csr4ad"), "dad": ("cbm.csr4dad", "csr4dad")}


def _class(table, product):
    import importlib
    mod, name = table[product]
    return getattr(importlib.import_module(mod), name)


def time_loop(fn, steps, warmup, stream):
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    for _ in range(steps):
        fn()
    stop.record(stream)
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def build_workload(name, alpha, cols, dev):
    from cbm import structure
    w = dict(WORKLOADS[name])
    if alpha is not None:
        w["alpha"] = alpha
    if cols is not None:
        w["cols"] = cols
    t0 = time.perf_counter()
    ei = structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"])
    log(f"graph {name}: N={w['n']} nnz={ei.shape[1]} ({time.perf_counter() - t0:.1f}s)")
    return _build_cbm(w, ei, dev)


def read_graph(path):
    """A local graph file: MatrixMarket (.mtx), numpy edge_index (.npy /
    .npz), or a whitespace-separated edge list (anything else)."""
    from cbm import datasets
    if path.endswith(".mtx"):
        return datasets.read_matrix_market(path)
    if path.endswith((".npy", ".npz")):
        return datasets.read_edge_index(path)
    return datasets.read_edge_list(path)


def file_workload(path, product, alpha, cols, dev):
    """A workload over a graph read from a local file, with the reference
    harness's defaults (alpha = 32, F = 512: scripts/matmul.sh:9-17,
    benchmark/benchmark_matmul_cuda.py:16)."""
    t0 = time.perf_counter()
    ei, n = read_graph(path)
    log(f"graph {path}: N={n} nnz={ei.shape[1]} ({time.perf_counter() - t0:.1f}s)")
    w = {"n": n, "alpha": 32 if alpha is None else alpha, "cols": 512 if cols is None else cols,
         "product": product or "mm"}
    return _build_cbm(w, ei, dev)


def _build_cbm(w, ei, dev):
    import torch
    t0 = time.perf_counter()
    m = _class(CBM_CLASSES, w["product"])(ei, torch.ones(ei.shape[1]), alpha=w["alpha"], device=dev)
    w["build_s"] = round(time.perf_counter() - t0, 2)
    h = m.cusparse_extension
    log(f"CBM built ({CBM_CLASSES[w['product']][1]}): nnz(Delta)={h.nnz} tree edges={h.n_edges} depth={h.depth} "
        f"({w['build_s']:.1f}s)")
    return w, ei, m


def build_stages(ei, n, alpha, dev, stream):
    """Times the stages of the structure build: the adjacency CSR (host), the
    candidate-edge scan on the GPU (HIP events around count + emit, inputs
    already on the device) and on the host cores, and the arborescence plus
    Delta (host).  Both scans must give the same list."""
    import torch
    from cbm import structure
    cores = len(os.sched_getaffinity(0))
    out = {"host_cores": cores}
    t0 = time.perf_counter()
    adj = structure.Adjacency(ei[0], ei[1], torch.ones(ei.shape[1]), n)
    out["adjacency_s"] = round(time.perf_counter() - t0, 3)
    structure.gpu_candidates(adj, alpha, dev)  # warm-up (first launch, allocator)
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    start.record(stream)
    gsrc, gdst, gw = structure.gpu_candidates(adj, alpha, dev)
    stop.record(stream)
    torch.cuda.synchronize()
    out["candidates_gpu_ms"] = round(start.elapsed_time(stop), 3)
    out["candidates_gpu_wall_ms"] = round(1e3 * (time.perf_counter() - t0), 3)
    t0 = time.perf_counter()
    hsrc, hdst, hw = structure.host_candidates(adj, alpha)
    out["candidates_host_s"] = round(time.perf_counter() - t0, 3)
    out["n_candidates"] = int(hsrc.numel())
    out["scans_identical"] = bool(torch.equal(gsrc.cpu(), hsrc) and torch.equal(gdst.cpu(), hdst)
                                  and torch.equal(gw.cpu(), hw))
    t0 = time.perf_counter()
    structure.build_from_candidates(adj, hsrc, hdst, hw)
    out["arborescence_delta_s"] = round(time.perf_counter() - t0, 3)
    adj.close()
    log(f"build stages: {out}")
    return out


def measure(name, w, ei, m, args, dev, stream):
    """Times the CBM step and its phases, the CSR comparators, and returns the
    record for one workload (plus the device tensors the CPU leg checks)."""
    import torch
    from cbm.cbm_cusparse import cusparseCBM

    h = m.cusparse_extension
    n, f = w["n"], w["cols"]
    nnz = int(ei.shape[1])
    gen = torch.Generator(device=dev).manual_seed(2005)
    x32 = torch.randn(n, f, device=dev, generator=gen)
    g32 = torch.randn(n, f, device=dev, generator=gen)
    x16, g16 = x32.to(torch.bfloat16), g32.to(torch.bfloat16)
    y = torch.empty(n, f, device=dev)
    dx = torch.empty(n, f, device=dev)
    z = torch.empty(n, f, device=dev)
    h.reserve(f)

    def step(xx, gg):
        def run():
            h.forward(xx, y)
            h.matmul_t(gg, dx, workspace=z)
        return run

    def graph_ms(fn):
        """The same step captured once in a HIP graph and replayed: GPU time
        without the per-call host work (Python checks, launches)."""
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            fn()
        torch.cuda.current_stream().wait_stream(side)
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            fn()
        ms = time_loop(graph.replay, K, W, stream)
        del graph
        return ms

    K, W = args.steps, args.warmup
    ms16 = time_loop(step(x16, g16), K, W, stream)
    ms32 = time_loop(step(x32, g32), K, W, stream)
    ms16_graph = graph_ms(step(x16, g16))
    ms32_graph = graph_ms(step(x32, g32))
    log(f"{name}: CBM step bf16 {ms16:.3f} ms, fp32 {ms32:.3f} ms; graph-replayed {ms16_graph:.3f} / {ms32_graph:.3f}")
    fused = {"bf16": h.fused_active(torch.bfloat16, f), "fp32": h.fused_active(torch.float32, f)}
    over_a = {"bf16": h.uncompressed_active(torch.bfloat16, f), "fp32": h.uncompressed_active(torch.float32, f)}
    # the same step in the other forms: compressed where the cost model runs
    # it over A, the one-launch forward (forced on where the plan has units),
    # and the adjoint through the Delta^T pass instead of the forward structure
    forms = {}
    if any(over_a.values()):
        h.attach_uncompressed(h._plain, False)
        forms["compressed"] = {"bf16": round(time_loop(step(x16, g16), K, W, stream), 4),
                               "fp32": round(time_loop(step(x32, g32), K, W, stream), 4)}
        h.attach_uncompressed(h._plain, None)
    if h.plan_stats.get("units", 0) > 0:
        h.set_fused(True)
        if h._plain is not None:
            h.attach_uncompressed(h._plain, False)
        if h.fused_active(torch.bfloat16, f):
            forms["one_launch_forward"] = {"bf16": round(time_loop(step(x16, g16), K, W, stream), 4),
                                           "fp32": round(time_loop(step(x32, g32), K, W, stream), 4)}
        if h._plain is not None:
            h.attach_uncompressed(h._plain, None)
        h.set_fused(None)
    if h.symmetric:
        h.set_symmetric(False)
        forms["delta_t_adjoint"] = {"bf16": round(time_loop(step(x16, g16), K, W, stream), 4),
                                    "fp32": round(time_loop(step(x32, g32), K, W, stream), 4)}
        h.set_symmetric(True)
    ordered = {"bf16": h.order_active(torch.bfloat16, f), "fp32": h.order_active(torch.float32, f)}
    if getattr(h, "_order", None) is not None or getattr(getattr(h, "_plain", None), "_order", None) is not None:
        h.set_order(False)
        forms["position_order"] = {"bf16": round(time_loop(step(x16, g16), K, W, stream), 4),
                                   "fp32": round(time_loop(step(x32, g32), K, W, stream), 4)}
        h.set_order(None)
    log(f"{name}: over A {over_a}; one launch {fused}; gather order {ordered}; other forms {forms}")
    ph = {
        "forward": time_loop(lambda: h.forward(x16, y), K, W, stream),
        "forward_delta_spmm": time_loop(lambda: h.delta_spmm(x16, y), K, W, stream),
        "forward_propagate": time_loop(lambda: h.propagate(y), K, W, stream),
        "adjoint": time_loop(lambda: h.matmul_t(g16, dx, workspace=z), K, W, stream),
        # the two halves of the Delta^T form of the adjoint
        "adjoint_reverse": time_loop(lambda: h.adjoint_reverse(g16, z), K, W, stream),
        "adjoint_delta_t": time_loop(lambda: h.adjoint_delta_t(g16, z, dx), K, W, stream),
    }
    launches = {}
    h.forward(x16, y)
    launches["forward"] = h.last_launches()
    h.matmul_t(g16, dx, workspace=z)
    launches["adjoint"] = h.last_launches()
    # the other multi-launch form of the forward (same bits): per-level Delta + parent passes
    h.set_fused(False)
    h.set_forward_split(False)
    ph["forward_fused_per_level"] = time_loop(lambda: h.forward(x16, y), K, W, stream)
    h.set_forward_split(None)
    h.set_fused(None)
    # feature slabs spread over the 8 XCDs (same bits)
    xcd = {}
    for slabs in (2, 4, 8):
        h.set_xcd_slabs(slabs)
        xcd[slabs] = {"step_bf16": round(time_loop(step(x16, g16), K, W, stream), 4),
                      "forward": round(time_loop(lambda: h.forward(x16, y), K, W, stream), 4),
                      "adjoint_delta_t": round(time_loop(lambda: h.adjoint_delta_t(g16, z, dx), K, W, stream), 4)}
    h.set_xcd_slabs(1)
    log(f"{name}: XCD slabs {xcd}")
    # gathers by XCD column range (different rounding order): off, and each piece set
    xr = {}
    if h.plan_stats.get("fwd_pieces", 0) > 0:
        for on, phases in ((False, 1), (8, 1), (4, 1), (8, 2), (4, 2)):
            h.set_xcd_ranges(on)
            h.set_xcd_phases(phases)
            xr[f"{on}_ranges_{phases}_phases" if on else "off"] = {
                "step_bf16": round(time_loop(step(x16, g16), K, W, stream), 4),
                "step_fp32": round(time_loop(step(x32, g32), K, W, stream), 4),
                "forward": round(time_loop(lambda: h.forward(x16, y), K, W, stream), 4),
                "adjoint_delta_t": round(time_loop(lambda: h.adjoint_delta_t(g16, z, dx), K, W, stream), 4)}
        h.set_xcd_ranges(None)
        h.set_xcd_phases(None)
        log(f"{name}: XCD ranges {xr}")
    # results the CPU leg compares against FP64 (fresh, from the BF16 and FP32 steps)
    out = {}
    for tag, (xx, gg) in (("bf16", (x16, g16)), ("fp32", (x32, g32))):
        yy, dd = torch.empty(n, f, device=dev), torch.empty(n, f, device=dev)
        h.forward(xx, yy)
        h.matmul_t(gg, dd, workspace=z)
        out[tag] = (yy, dd)

    # the BF16 results against the FP32 CBM results (same kernels, FP32 inputs)
    vs_fp32 = {k: float(((out["bf16"][i] - out["fp32"][i]).abs().max() / out["fp32"][i].abs().max()).item())
               for i, k in enumerate(("forward", "adjoint"))}

    # rocSPARSE CSR FP32 of the same product M: the fastest algorithm on the
    # forward, M^T as an explicit CSR
    csr = _class(CSR_CLASSES, w["product"])(ei, num_nodes=n, device=dev)
    ext = csr.cusparse_extension
    ext.transposed()
    alg_ms = {}
    for alg_name, alg in ROCSPARSE_ALGS.items():
        try:
            ext.set_algorithm(alg)
            alg_ms[alg_name] = time_loop(lambda: csr.matmul(x32, y), max(3, K // 4), 2, stream)
        except Exception as exc:  # an algorithm rocSPARSE rejects for this shape
            log(f"rocSPARSE {alg_name}: {exc}")
    best_alg = min(alg_ms, key=alg_ms.get)
    ext.set_algorithm(ROCSPARSE_ALGS[best_alg])

    def step_csr():
        csr.matmul(x32, y)
        csr.matmul_t(g32, dx)

    ms_csr = time_loop(step_csr, K, W, stream)
    log(f"{name}: rocSPARSE CSR fp32 step {ms_csr:.3f} ms ({best_alg})")
    m_csr = (csr.a.to("cpu"), ext.transposed().a.to("cpu"))  # M and M^T (FP32 values) for the host checks
    del csr, ext

    # the same gathers without compression: our kernels over plain A (a tree
    # with no edges) with M's scalings, no Delta
    a = torch.sparse_csr_tensor(*_csr_of(ei, n), (n, n)).to(dev)
    e0 = torch.zeros(0, dtype=torch.int32, device=dev)
    hip_csr = cusparseCBM(a, torch.zeros(1, dtype=torch.int32, device=dev), e0, e0, torch.empty(0, device=dev))
    if h.mode != hip_csr.mode:
        hip_csr = hip_csr.with_scaling(h.mode, h._d)
    hip_csr.reserve(f)
    ms_hip_csr16 = time_loop(lambda: (hip_csr.forward(x16, y), hip_csr.matmul_t(g16, dx, workspace=z)), K, W, stream)
    ms_hip_csr32 = time_loop(lambda: (hip_csr.forward(x32, y), hip_csr.matmul_t(g32, dx, workspace=z)), K, W, stream)
    del hip_csr, a

    n_internal = int(h._dev["rev_node"].numel())
    scaled = w["product"] != "mm"
    b16 = phase_bytes(n, h.nnz, h.n_edges, n_internal, h.depth, f, 2, scaled, h.symmetric)
    b32 = phase_bytes(n, h.nnz, h.n_edges, n_internal, h.depth, f, 4, scaled, h.symmetric)
    # The step's largest single launch: the one-launch forward where it runs,
    # else the forward's Delta pass (chunks and short rows in one grid); the
    # adjoint of a symmetric A is the same kernel over G.
    dom = "forward" if fused["bf16"] else "forward_delta_spmm"
    dom_gbs = b16[dom] / (ph[dom] * 1e-3) / 1e9
    gen_desc = (f"copying model N={n} blocks={w['blocks']} p={w['p']} lambda={w['lam']} seed_degree={w['seed_deg']} "
                f"seed={w['seed']}" if "blocks" in w else f"graph file N={n}")
    rec = {
        "workload": f"{name}: {CBM_CLASSES[w['product']][1]} fwd+bwd, {gen_desc}, nnz(A)={nnz}, alpha={w['alpha']}, "
                    f"F={f}",
        "nnz_delta": h.nnz, "tree_edges": h.n_edges, "internal_nodes": n_internal, "depth": h.depth,
        "compression": round(nnz / max(1, h.nnz), 3),
        "ms_step_bf16": round(ms16, 4), "ms_step_fp32": round(ms32, 4),
        "ms_step_bf16_graph_replay": round(ms16_graph, 4), "ms_step_fp32_graph_replay": round(ms32_graph, 4),
        "ms_phase_bf16": {k: round(v, 4) for k, v in ph.items()},
        "ms_step_rocsparse_csr_fp32": round(ms_csr, 4), "rocsparse_alg": best_alg,
        "ms_rocsparse_fwd_by_alg": {k: round(v, 4) for k, v in alg_ms.items()},
        "ms_step_hip_csr_bf16": round(ms_hip_csr16, 4), "ms_step_hip_csr_fp32": round(ms_hip_csr32, 4),
        "speedup_bf16_vs_rocsparse_fp32": round(ms_csr / ms16, 3),
        "speedup_fp32_vs_rocsparse_fp32": round(ms_csr / ms32, 3),
        "gbs_step_bf16": round(b16["step"] / (ms16 * 1e-3) / 1e9, 1),
        "gbs_step_fp32": round(b32["step"] / (ms32 * 1e-3) / 1e9, 1),
        "gbs_rocsparse_csr": round(2 * csr_bytes(n, nnz, f) / (ms_csr * 1e-3) / 1e9, 1),
        "bytes_step_bf16": b16["step"], "bytes_phase_bf16": b16,
        "dominant_phase": dom, "dominant_gbs": round(dom_gbs, 1),
        "gather_gbs_dominant": round(b16["gather_delta"] / (ph[dom] * 1e-3) / 1e9, 1),
        "symmetric_adjoint": h.symmetric, "one_launch_forward": fused, "runs_over_a": over_a, "launches": launches,
        "gather_order": ordered,
        "ms_step_other_forms": forms,
        "max_rel_error_bf16_vs_fp32_cbm": vs_fp32,
        "ms_xcd_slabs_bf16": xcd,
        "ms_xcd_ranges": xr,
        "xcd_ranges_default": {"bf16": h.xcd_ranges_active(torch.bfloat16, f),
                               "fp32": h.xcd_ranges_active(torch.float32, f)},
    }
    return rec, out, (x32, g32, x16, g16), m_csr


def gcn_step(args, dev, stream):
    """Training step of a 2-layer GCN over D^-1/2 A D^-1/2 (hidden 256, FP32)
    on the c3 graph (N = 169,343, alpha = 16): forward, loss <out, G>, both
    adjoints, SGD update.  CBM kernels vs the rocSPARSE CSR product."""
    import torch
    from cbm import structure
    from cbm.cbm4dad import cbm4dad
    from cbm.csr4dad import csr4dad
    from cbm.gcn import GCN
    w = WORKLOADS["c3"]
    n = w["n"]
    ei = structure.copying_model(n, w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"])
    gen = torch.Generator(device=dev).manual_seed(2003)
    x = torch.randn(n, 128, device=dev, generator=gen)
    g = torch.randn(n, 256, device=dev, generator=gen)
    res = {}
    for tag, adj in (("cbm", cbm4dad(ei, torch.ones(ei.shape[1]), alpha=w["alpha"], device=dev)),
                     ("rocsparse_csr", csr4dad(ei, num_nodes=n, device=dev))):
        model = GCN(adj, 128, 256, 256, seed=5, device=dev)
        opt = torch.optim.SGD(model.parameters(), lr=1e-4)

        def step():
            opt.zero_grad(set_to_none=False)
            loss = (model(x) * g).sum()
            loss.backward()
            opt.step()

        res[tag] = time_loop(step, args.steps, args.warmup, stream)
        log(f"gcn {tag}: {res[tag]:.3f} ms per training step")
    out = {"metric": "2-layer GCN training step ms (cbm4dad fwd + adjoint, FP32)", "value": round(res["cbm"], 4),
           "unit": "ms", "higher_is_better": False, "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
           "config": {"workload": f"c3: N={n}, nnz(A)={int(ei.shape[1])}, alpha={w['alpha']}, X Nx128, G Nx256, "
                                  "hidden 256"},
           "detail": {"ms_step_cbm": round(res["cbm"], 4), "ms_step_rocsparse_csr": round(res["rocsparse_csr"], 4),
                      "speedup": round(res["rocsparse_csr"] / res["cbm"], 3)}}
    print(json.dumps(out), flush=True)


def gather_ceiling(name, w, m, args, dev, stream):
    """The ceiling of the gather passes: the same Delta-pass kernel over
    uniform-random column indices of the workload's shape (N rows, the mean
    entries per row of its Delta, X in BF16 at its F), columns drawn over all
    N rows (X beyond the 256 MB Infinity Cache at c5) and over a
    400,000-row span (about 100 MB of X at F = 128, inside it).  Reported in
    gathered GB/s (entries x 2 F bytes per call) beside the same kernel over
    the workload's own Delta."""
    import torch
    from cbm._hip_sigs import MODE_AX
    from cbm.cbm_cusparse import cusparseCBM
    K, W = args.steps, args.warmup
    n, f = w["n"], w["cols"]
    h0 = m.cusparse_extension
    k = max(1, round(h0.nnz / n))
    gen = torch.Generator(device=dev).manual_seed(2005)
    x16 = torch.randn(n, f, device=dev, generator=gen).to(torch.bfloat16)
    y = torch.empty(n, f, device=dev)
    h0.reserve(f)
    ms_real = time_loop(lambda: h0.delta_spmm(x16, y), K, W, stream)
    out = {"sample": f"N={n}, {k} uniform-random entries per row, F={f} BF16",
           "workload_delta_pass": {"ms": round(ms_real, 4), "entries": h0.nnz,
                                   "gather_gbs": round(h0.nnz * 2 * f / (ms_real * 1e-3) / 1e9, 1)}}
    e0 = torch.zeros(0, dtype=torch.int32, device=dev)
    for tag, span in (("above_mall", n), ("inside_mall", min(n, 400_000))):
        cpu_gen = torch.Generator().manual_seed(17)
        col = torch.randint(0, span, (n, k), generator=cpu_gen, dtype=torch.int32).sort(dim=1).values.reshape(-1)
        crow = torch.arange(0, n * k + 1, k, dtype=torch.int32)
        a = torch.sparse_csr_tensor(crow, col, torch.ones(n * k), (n, n)).to(dev)
        del col, crow
        h = cusparseCBM(a, torch.zeros(1, dtype=torch.int32, device=dev), e0, e0, torch.empty(0, device=dev),
                        mode=MODE_AX)
        h.reserve(f)
        ms = time_loop(lambda: h.delta_spmm(x16, y), K, W, stream)
        out[tag] = {"span_rows": span, "x_bytes_gathered_from": span * f * 2, "ms": round(ms, 4),
                    "gather_gbs": round(n * k * 2 * f / (ms * 1e-3) / 1e9, 1)}
        log(f"gather ceiling {tag}: {out[tag]}")
        del h, a
        torch.cuda.empty_cache()
    out["workload_frac_of_ceiling"] = round(out["workload_delta_pass"]["gather_gbs"] / out["above_mall"]["gather_gbs"],
                                            3)
    print(json.dumps({name: out}), flush=True)


def profile_steps(w, m, args, dev, stream):
    """The CBM step alone (BF16 then FP32 storage), for kernel traces and
    counter runs: args.steps timed steps after args.warmup, per dtype; then,
    per dtype, the Delta pass alone (cbm_delta_spmm over X) timed the way the
    bench times its phases (args.steps calls after args.warmup), so that the
    trace's launches of that loop can be set beside the bench's figure."""
    import torch
    h = m.cusparse_extension
    n, f = w["n"], w["cols"]
    gen = torch.Generator(device=dev).manual_seed(2005)
    x32 = torch.randn(n, f, device=dev, generator=gen)
    g32 = torch.randn(n, f, device=dev, generator=gen)
    y, dx, z = (torch.empty(n, f, device=dev) for _ in range(3))
    h.reserve(f)
    inputs = (("bf16", x32.to(torch.bfloat16), g32.to(torch.bfloat16)), ("fp32", x32, g32))
    for tag, xx, gg in inputs:
        ms = time_loop(lambda: (h.forward(xx, y), h.matmul_t(gg, dx, workspace=z)), args.steps, args.warmup, stream)
        log(f"profile {tag}: {ms:.3f} ms per step ({args.steps} steps)")
    for tag, xx, _ in inputs:
        ms = time_loop(lambda: h.delta_spmm(xx, y), args.steps, args.warmup, stream)
        log(f"profile {tag}: Delta pass alone {ms:.3f} ms per call ({args.steps} calls after {args.warmup})")


def _csr_of(ei, n):
    import torch
    crow = torch.zeros(n + 1, dtype=torch.int64)
    crow[1:] = torch.bincount(ei[0].long(), minlength=n).cumsum(0)
    return crow.to(torch.int32), ei[1].to(torch.int32), torch.ones(ei.shape[1])


def _host_csr(t, dtype):
    """A CPU CSR tensor with int64 indices and values of `dtype`."""
    import torch
    return torch.sparse_csr_tensor(t.crow_indices().to(torch.int64), t.col_indices().to(torch.int64),
                                   t.values().to(dtype), t.shape)


def cpu_csr_mkl(m_csr, n, f, inputs, budget_s=20.0):
    """torch.sparse CSR FP32 M @ X and M^T @ G (MKL) on the host cores, on the
    full graph, for as many steps as fit in the budget (at most 10): the
    reference harness's CPU CSR product (benchmark/compare_matmul_cuda.py:53),
    reported beside the oracle's time."""
    import torch
    cores = len(os.sched_getaffinity(0))
    torch.set_num_threads(cores)
    a, at = _host_csr(m_csr[0], torch.float32), _host_csr(m_csr[1], torch.float32)
    x32, g32 = inputs[0].cpu(), inputs[1].cpu()
    _ = a @ x32[:, :8]  # warm MKL
    times = []
    t_end = time.perf_counter() + budget_s
    while True:
        t0 = time.perf_counter()
        y = a @ x32
        dx = at @ g32
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 10:
            break
    ms = 1e3 * sum(times) / len(times)
    del y, dx
    return {"value": round(ms, 3), "unit": "ms per M@X fwd+bwd (FP32)", "cores": cores,
            "kind": "torch.sparse CSR (MKL)",
            "sample": f"full graph N={n} F={f}, {len(times)} steps of torch.sparse CSR M@X + M^T@G (MKL, FP32)"}


def cpu_baseline_oracle(m, product, n, f, x32, g32, budget_s=20.0, cols=4):
    """The oracle's CBM product timed on the host: Delta' X in FP64 (scipy,
    one thread) followed by the reference's branch-serial propagation
    (cbm/cbm.cu:25-91) restated in C (oracle/cbm_oracle.c, one thread), for
    the forward and for the adjoint (for a symmetric A the same form over G;
    D A for the A D product).  A bounded sample: `cols` columns of X and G on
    the full graph, repeated for about budget_s seconds; the time is scaled
    to the full F (the work is linear in the columns)."""
    import ctypes
    import numpy as np
    import scipy.sparse as sp
    import torch
    from oracle import oracle as O
    crow = m.deltas.crow_indices().cpu().numpy().astype(np.int64)
    col = m.deltas.col_indices().cpu().numpy()
    sign = np.sign(m.deltas.values().cpu().numpy()).astype(np.float64)
    delta = sp.csr_matrix((sign, col, crow), shape=(n, n))
    mode = {"mm": "ax", "ad": "adx", "dad": "dadx"}[product]
    dvec = None
    if mode != "ax":
        dvec = m.D.cpu().numpy().astype(np.float64)  # D^-1/2 as the product uses it
        delta_f = delta @ sp.diags(dvec)
        delta_f = (sp.diags(dvec) @ delta_f).tocsr() if mode == "dadx" else delta_f.tocsr()
        delta_t = (sp.diags(dvec) @ delta).tocsr()  # D Delta: the adjoint of A D is D A
    else:
        delta_f = delta_t = delta
    br = np.ascontiguousarray(m.mca_offset.cpu().numpy(), np.int32)
    src = np.ascontiguousarray(m.mca_src_idx.cpu().numpy(), np.int32)
    dst = np.ascontiguousarray(m.mca_dst_idx.cpu().numpy(), np.int32)
    xs = np.ascontiguousarray(x32[:, :cols].cpu().numpy(), np.float64)
    gs = np.ascontiguousarray(g32[:, :cols].cpu().numpy(), np.float64)
    mult_f = dvec.ctypes.data_as(ctypes.c_void_p) if mode == "dadx" else None
    mult_t = dvec.ctypes.data_as(ctypes.c_void_p) if mode in ("adx", "dadx") else None
    lib = O.lib()

    def product(dm, v, mult):
        y = np.ascontiguousarray(dm @ v)
        lib.oracle_update(cols, y, len(br) - 1, br, src, dst, mult)
        return y

    times = []
    t_end = time.perf_counter() + budget_s
    while True:
        t0 = time.perf_counter()
        product(delta_f, xs, mult_f)
        product(delta_t, gs, mult_t)
        times.append(time.perf_counter() - t0)
        if time.perf_counter() > t_end or len(times) >= 20:
            break
    ms = 1e3 * sum(times) / len(times) * (f / cols)
    return {"value": round(ms, 3), "unit": "ms per M@X fwd+bwd (FP64 oracle, scaled to F)", "cores": 1,
            "kind": "port",
            "sample": f"full graph N={n}, {cols} of F={f} columns, {len(times)} steps (forward + adjoint) of the "
                      f"oracle's CBM product: scipy Delta' X (FP64) + the branch-serial propagation in C, "
                      f"{1e3 * sum(times) / len(times):.1f} ms per sampled step, scaled by F/{cols}"}


def fp64_errors(m_csr, gpu_out, inputs):
    """Max relative error (max-norm) of the GPU results against FP64 products
    on the host: FP32 inputs, BF16-rounded inputs, and the BF16 results
    against FP64 of the original FP32 inputs (end to end)."""
    import torch
    a64, at64 = _host_csr(m_csr[0], torch.float64), _host_csr(m_csr[1], torch.float64)
    x32, g32 = inputs[0].cpu(), inputs[1].cpu()

    def rel(got, ref):
        return float((got.double().cpu() - ref).abs().max() / ref.abs().max())

    errors = {}
    for tag, xx, gg in (("fp32", x32, g32), ("bf16", inputs[2].float().cpu(), inputs[3].float().cpu())):
        ry, rd = a64 @ xx.double(), at64 @ gg.double()
        errors[tag] = {"forward": rel(gpu_out[tag][0], ry), "adjoint": rel(gpu_out[tag][1], rd)}
    ry, rd = a64 @ x32.double(), at64 @ g32.double()
    errors["bf16_vs_fp32_inputs"] = {"forward": rel(gpu_out["bf16"][0], ry), "adjoint": rel(gpu_out["bf16"][1], rd)}
    return errors


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c5", choices=sorted(WORKLOADS))
    ap.add_argument("--alpha", type=int, default=None)
    ap.add_argument("--cols", type=int, default=None)
    ap.add_argument("--secondary", default=None,
                    help="second workload reported under 'secondary' ('' for none; default c4, none with --graph)")
    ap.add_argument("--graph", default=None, metavar="FILE",
                    help="measure a local graph file (MatrixMarket .mtx, numpy edge_index .npy/.npz, or an edge "
                         "list) instead of a synthetic workload; alpha defaults to 32 and F to 512 there")
    ap.add_argument("--product", default=None, choices=sorted(CBM_CLASSES),
                    help="the product measured on a --graph file (default mm)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--gcn", action="store_true",
                    help="time the 2-layer GCN training step (c3 graph, hidden 256) instead; CBM vs rocSPARSE CSR")
    ap.add_argument("--profile", action="store_true",
                    help="run only the CBM steps (BF16, then FP32) for a profiler; prints their times")
    ap.add_argument("--build-only", action="store_true",
                    help="build the workload's structure and time the build stages only (GPU vs host candidate scan)")
    ap.add_argument("--gather-ceiling", action="store_true",
                    help="time the Delta-pass kernel over uniform-random indices of the workload's shape (X beyond "
                         "and inside the Infinity Cache) beside its own Delta pass; prints one JSON line")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="first run one plain device copy of MIB MiB (the calibration kernel of counter runs)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: the CBM product here runs on a single MI355X; "
              "the multi-GPU sharding is a design note only. Run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import torch
    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream()
    if args.copy_probe > 0:
        # A copy of known size: its FETCH_SIZE / WRITE_SIZE per dispatch give the
        # counters' scale against bytes (read once, written once).
        n_floats = args.copy_probe * (1 << 20) // 4
        src = torch.ones(n_floats, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"copy probe: {args.copy_probe} MiB read + {args.copy_probe} MiB written")
        del src, dst

    if args.gcn:
        gcn_step(args, dev, stream)
        return
    name = os.path.basename(args.graph) if args.graph else args.workload
    if args.secondary is None:
        args.secondary = "" if args.graph else "c4"
    if args.build_only:
        from cbm import structure
        if args.graph:
            ei, n = read_graph(args.graph)
            w = {"n": n, "alpha": 32 if args.alpha is None else args.alpha}
        else:
            w = dict(WORKLOADS[args.workload])
            if args.alpha is not None:
                w["alpha"] = args.alpha
            ei = structure.copying_model(w["n"], w["blocks"], w["p"], w["lam"], w["seed_deg"], w["seed"])
        stages = build_stages(ei, w["n"], w["alpha"], dev, stream)
        print(json.dumps({"workload": name, "n": w["n"], "nnz": int(ei.shape[1]), "alpha": w["alpha"],
                          "build_stages": stages}), flush=True)
        return
    if args.graph:
        w, ei, m = file_workload(args.graph, args.product, args.alpha, args.cols, dev)
    else:
        w, ei, m = build_workload(args.workload, args.alpha, args.cols, dev)
    if args.profile:
        profile_steps(w, m, args, dev, stream)
        return
    if args.gather_ceiling:
        gather_ceiling(name, w, m, args, dev, stream)
        return
    rec, gpu_out, inputs, m_csr = measure(name, w, ei, m, args, dev, stream)
    rec["build_s"] = w["build_s"]
    rec["build_stages"] = build_stages(ei, w["n"], w["alpha"], dev, stream)
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline_oracle(m, w["product"], w["n"], w["cols"], inputs[0], inputs[1])
        log(f"CPU baseline (oracle) {cpu['value']} ms on {cpu['cores']} core ({time.perf_counter() - t0:.1f}s)")
        t0 = time.perf_counter()
        rec["cpu_csr_mkl"] = cpu_csr_mkl(m_csr, w["n"], w["cols"], inputs)
        log(f"CPU torch.sparse CSR {rec['cpu_csr_mkl']['value']} ms on {rec['cpu_csr_mkl']['cores']} cores "
            f"({time.perf_counter() - t0:.1f}s)")
        rec["max_rel_error_vs_fp64"] = fp64_errors(m_csr, gpu_out, inputs)
    del gpu_out, inputs, m, m_csr

    secondary = None
    if args.secondary and args.secondary != name:
        w2, ei2, m2 = build_workload(args.secondary, None, None, dev)
        secondary, out2, in2, m_csr2 = measure(args.secondary, w2, ei2, m2, args, dev, stream)
        secondary["build_s"] = w2["build_s"]
        secondary["build_stages"] = build_stages(ei2, w2["n"], w2["alpha"], dev, stream)
        if not args.no_cpu:
            secondary["max_rel_error_vs_fp64"] = fp64_errors(m_csr2, out2, in2)
        del m2, out2, in2, m_csr2

    ms16 = rec["ms_step_bf16"]
    dom = rec["dominant_phase"]
    traffic = None if args.graph else measured_traffic(args.workload, dom)
    ceiling = None if args.graph else committed_gather_ceiling(args.workload, rec)
    out = {
        "metric": "A@X fwd+bwd ms & HBM GB/s (% roofline), CBM-HIP vs CSR on-GPU, FP32 & BF16",
        "value": ms16,
        "unit": "ms per fwd+bwd step (BF16 storage, FP32 accumulate)",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms16,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16-storage/f32-accumulate",
        "data": (f"graph file {name}; X, G ~ N(0,1)" if args.graph
                 else "synthetic (seeded community copying model; X, G ~ N(0,1))"),
        "config": {"workload": rec["workload"], "nnz_delta": rec["nnz_delta"], "depth": rec["depth"],
                   "compression": rec["compression"]},
        "roofline": {"bound": "hbm", "achieved": rec["dominant_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(rec["dominant_gbs"] / HBM_PEAK_GBS, 4),
                     # L2-to-fabric bytes per call (FETCH_SIZE + WRITE_SIZE): they include
                     # Infinity Cache hits, and no counter on this system separates the bytes
                     # that came from HBM; committed from a counter run, not measured here
                     "traffic": traffic.get("bytes") if traffic else None,
                     "traffic_kind": "L2-to-fabric bytes (Infinity Cache hits included); HBM-only bytes unmeasurable",
                     "traffic_hbm": None,
                     "traffic_gbs": round(traffic["bytes"] / (rec["ms_phase_bf16"][dom] * 1e-3) / 1e9, 1)
                     if traffic else None,
                     "traffic_source": (f"committed {traffic.get('source')}" if traffic else None),
                     "kernel": (f"{dom} (BF16): " + ("k_fused_forward (one launch: Delta pieces + LDS propagation)"
                                                     if dom == "forward" else
                                                     "the Delta pass (k_row_mixed: long-row chunks + short rows, "
                                                     "in the gather order where the matrix has one, or "
                                                     "k_row_pass over XCD range pieces) + k_row_reduce")),
                     "algorithmic_bytes": rec["bytes_phase_bf16"][dom],
                     "ms": rec["ms_phase_bf16"][dom],
                     "gather_bytes": rec["bytes_phase_bf16"]["gather_delta"],
                     "gather_gbs": rec["gather_gbs_dominant"],
                     "gather_ceiling_gbs": ceiling["gbs"] if ceiling else None,
                     "frac_of_gather_ceiling": round(rec["gather_gbs_dominant"] / ceiling["gbs"], 3)
                     if ceiling else None,
                     "gather_ceiling_source": ceiling["source"] if ceiling else None},
        "detail": rec,
        "cpu_baseline": cpu,
        "secondary": secondary,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
