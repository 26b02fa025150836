"""Throughput of the TBL fractional-step time step on one MI355X.

One "step" is the reference's whole time step (dnssetbl/dnssetbl.f90:253-1262):
recycling inflow, convection + pressure gradient, implicit diffusion
(Gauss-Seidel), PPE right-hand side, pressure-Poisson GMRES, projection,
boundary conditions, divergence / mass-flux diagnostics, statistics and the
kinetic-energy guard, on fields resident in HBM.

Default workload: the 1024x192x256 case (reference axes nX=1024 streamwise,
nY=256 span, nZ=192 wall-normal; N = 50,331,648 cells) with its physical
set-up (turbulentbl_amd/cases.py): Re_theta,in = 1000 (Re_tau = 422.3,
U_inf/u_tau = 22.12), domain 60 x 3 x 6 delta99 (stream x wall x span),
tanh stretch 2.5, dt from CFL 0.3 on the initial field; seeded log-law +
divergence-free Fourier-mode field, recycling active (restart state),
GMRES(40) to a true relative residual of 1e-10 -- in mixed precision (FP32
fields/stencils/Krylov basis inside FP64 refinement) for `value`, and in
plain FP64 beside it.  Prints one JSON line.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mcell-steps/s mixed vs FP64 (div<=1e-8); %HBM peak; adjoint/forward cost"
HBM_PEAK_GBS = 8000.0

def log(*a):
    print(*a, file=sys.stderr, flush=True)


def make_inputs(case):
    """mesh (from the library), physics (Re, Ue from Re_theta; dt from CFL 0.3),
    initial field and recycling state of a configuration"""
    from turbulentbl_amd import cases, tbl
    c = cases.CASES[case]
    nX, nY, nZ = c["grid"]
    Re, Ue = cases.flow_parameters(case)
    # the mesh does not depend on dt: a probe context with a placeholder dt
    probe = tbl.Solver(nX, nY, nZ, **cases.solver_params(case, 1e-4, Re, Ue, kmax=1))
    mesh = probe.mesh()
    probe.close()
    t0 = time.time()
    Ut, rs = cases.initial_state(case, mesh, Re, Ue)
    dt = cases.cfl_dt(Ut, mesh)
    i_rec = cases.rec_plane(nX)
    uf = np.zeros((nY, nZ, 3), order="F")
    for p in range(3):
        uf[:, :, p] = Ut[i_rec - 1, :, :, p] - rs["avg"][i_rec - 1, None, :, p]
    phys = cases.describe(case, mesh, Re, Ue, dt)
    log("inputs: %s built in %.1fs; %s" % (c["name"], time.time() - t0, json.dumps(phys)))
    return dict(case=case, Ut=Ut, rs=rs, uf=uf, i_rec=i_rec, mesh=mesh, Re=Re, Ue=Ue, dt=dt, phys=phys)


HELM_MODES = {"zline": 1, "zebra": 3}
HELM = {"mode": "zebra"}                       # set from --helm


def make_solver(inp, precision, kmax, errtol, inner_tol, pc_extrap=0, **over):
    """A context at the workload's start state: seeded field, zero pressure,
    recycling statistics of the mean profile, step counter 1000 (recycling on)"""
    from turbulentbl_amd import cases, tbl
    nX, nY, nZ = cases.CASES[inp["case"]]["grid"]
    kw = cases.solver_params(inp["case"], inp["dt"], inp["Re"], inp["Ue"], precision=precision, kmax=kmax,
                             errtol=errtol, tol_mode=tbl.TBL_TOL_TRUE, inner_tol=inner_tol, mmax=200,
                             pc_extrap=pc_extrap, helm_mode=HELM_MODES[HELM["mode"]])
    kw.update(over)
    sv = tbl.Solver(nX, nY, nZ, **kw)
    rs = inp["rs"]
    sv.upload(inp["Ut"], np.zeros((nX, nY, nZ), order="F"), np.zeros((nX, nY, nZ), order="F"))
    sv.set_state(avg=rs["avg"], Wt=rs["Wt"], newk=rs["newk"], Ume=rs["Ume"], Ufluc_rec=inp["uf"],
                 bltkns=rs["bltkns"], lamda=1.0, lamda03=1.0, nstep=1000, cnt=0)
    return sv


def run_long_compare(inp, steps, every, kmax, errtol):
    """Mixed against FP64 over a long horizon: both contexts start from the
    same state and step in lockstep; every `every` steps compare the mean and
    per-station wall shear (C_f), the velocity (max norm) and the reference's
    divergence metric.  One JSON line per checkpoint on stderr; returns the
    list of checkpoints."""
    f64 = make_solver(inp, "fp64", kmax, errtol, 1e-12)
    mix = make_solver(inp, "mixed", kmax, errtol, 1e-4)
    rows = []
    for s in range(every, steps + 1, every):
        df, dm = f64.step(every), mix.step(every)
        tf, tm = f64.wall_shear()[1:-1], mix.wall_shear()[1:-1]
        uf, um = f64.download()[0], mix.download()[0]
        rows.append(dict(step=s, cf_mean_rel=float(abs(tm.mean() - tf.mean()) / abs(tf.mean())),
                         cf_station_max_rel=float(np.max(np.abs(tm - tf)) / np.max(np.abs(tf))),
                         u_max_rel=float(np.max(np.abs(um - uf)) / np.max(np.abs(uf))),
                         div_ref_metric_fp64=df[-1]["div"], div_ref_metric_mixed=dm[-1]["div"]))
        log(json.dumps(rows[-1]))
    f64.close()
    mix.close()
    return rows


def run_gpu(inp, precision, steps, warmup, kmax, errtol, inner_tol, fields=False, pc_extrap=0):
    """Warm-up steps, then `steps` timed steps (the throughput).  Inside the
    timed region only the z-line sweep class (the roofline kernel) carries HIP
    events, one pair per sweep around its two back-to-back passes, so that
    event records between the other launches do not enter the rate; a further
    `steps` steps with every class timed give the per-class breakdown, outside
    the timed region."""
    import torch
    from turbulentbl_amd import cases, tbl
    nX, nY, nZ = cases.CASES[inp["case"]]["grid"]
    N = nX * nY * nZ
    sv = make_solver(inp, precision, kmax, errtol, inner_tol, pc_extrap)
    log("%s: device bytes %.2f GB; warmup %d" % (precision, sv.device_bytes / 1e9, warmup))
    for w in range(warmup):
        d = sv.step(1)[0]
        log("  warmup %d: gs %s gmres %d cyc %d it  true_res %.3e" % (w, d["gs_sweeps"], d["gm_cycles"],
                                                                   d["gm_iters"], d["gm_true_res"]))
    sv.timer_enable(ROOF_CLASSES)
    sv.timer_read()
    torch.cuda.synchronize()
    sv.sync()
    t0 = time.perf_counter()
    diags = sv.step(steps)                     # one library call steps the K timed steps
    sv.sync()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    for s, d in enumerate(diags):
        log("  step %d: gs %s gmres %d cyc %d it  true_res %.3e  div %.3e" % (
            s, d["gs_sweeps"], d["gm_cycles"], d["gm_iters"], d["gm_true_res"], d["div"]))
    tm_roof = sv.timer_read()
    tau = sv.wall_shear()
    Ut = Press = None
    if fields:
        Ut, Press, _ = sv.download()
    # per-class breakdown (not timed for the rate)
    sv.timer_enable(tbl.KCLASSES)
    sv.timer_read()
    sv.sync()
    b0 = time.perf_counter()
    sv.step(steps)
    sv.sync()
    b1 = time.perf_counter()
    tm = sv.timer_read()
    step_dt = sv.p.dt
    sv.close()
    dt = t1 - t0
    return dict(rate=N * steps / dt / 1e6, ms_per_step=dt / steps * 1e3, diags=diags, timers=tm, tau=tau,
                timers_roof=tm_roof, ms_per_step_all_timed=(b1 - b0) / steps * 1e3,
                Ut=Ut, Press=Press, dt=step_dt, nint=(nX - 2) * (nY - 2) * (nZ - 2))


def run_reference_orderings(inp, P, steps=2, warmup=1):
    """The reference's own algorithm on the GPU, in FP64: per-slab
    lexicographic Gauss-Seidel (P slabs, halo planes from before each sweep),
    GMRES(50) with per-slab block SSOR and the preconditioned-residual errtol
    1e-5 rule; the same start state and P as the CPU baseline (the oracle runs
    this algorithm), one warm-up step from PCorr = 0, then `steps` timed
    steps, then the same steps again with every class timed."""
    from turbulentbl_amd import cases, tbl
    nX, nY, nZ = cases.CASES[inp["case"]]["grid"]
    N = nX * nY * nZ
    sv = make_solver(inp, "fp64", 50, 1e-5, 1e-3, helm_mode=tbl.TBL_HELM_LEX_SLAB,
                     precond=tbl.TBL_PC_SSOR_BLOCK, slabs=P, tol_mode=tbl.TBL_TOL_REF, mmax=1000)
    for _ in range(warmup):
        d = sv.step(1)[0]
        log("reference orderings warmup: gs %s gmres %d cyc %d it" % (d["gs_sweeps"], d["gm_cycles"], d["gm_iters"]))
    sv.sync()
    t0 = time.perf_counter()
    diags = sv.step(steps)
    sv.sync()
    dt = time.perf_counter() - t0
    sv.timer_enable(tbl.KCLASSES)
    sv.timer_read()
    sv.step(steps)
    sv.sync()
    tm = sv.timer_read()
    sv.close()
    per = {k: dict(count=v["count"], ms=round(v["ms"], 3)) for k, v in tm.items() if v["count"]}
    return dict(workload="the reference's solvers and orderings, FP64, %s on %d wall-normal slabs: per-slab "
                         "lexicographic Gauss-Seidel, GMRES(50) with per-slab block SSOR, errtol 1e-5 "
                         "(preconditioned residual)" % (cases.CASES[inp["case"]]["name"], P),
                slabs=P, value=round(N * steps / dt / 1e6, 4), unit="Mcell-steps/s",
                ms_per_step=round(dt / steps * 1e3, 3), gs_sweeps=[d["gs_sweeps"] for d in diags],
                gmres_iters=[d["gm_iters"] for d in diags], kernels=per)


# the class timed inside the timed region: the z-line Gauss-Seidel sweep,
# the dominant HBM-bound kernel of both precisions
ROOF_CLASSES = ("helm",)

# timer classes that are one HBM-bound kernel launch each (the preconditioner
# class is a chain of GEMMs and a line solve, the diagnostics class several
# small kernels: neither is a single kernel)
SINGLE_KERNEL = {"rhs": "k_momentum_rhs", "helm": "k_helm_zline", "helm_res": "k_helm_resid", "div": "k_ppe_rhs",
                 "poisson": "k_poisson", "dots": "k_dots", "update": "k_update", "resid": "k_resid",
                 "project": "k_project"}


def measured_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary (copy-calibrated); None if that kernel was not counted"""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_%s.txt" % workload)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = dict(w.split("=", 1) for w in line.split() if "=" in w)
            if parts.get("kernel") == kernel and "traffic_bytes" in parts:
                return float(parts["traffic_bytes"])
    return None


def roofline(tm, storage, workload):
    """dominant single kernel by total time; achieved = algorithmic bytes per
    launch / average launch time (HIP events on the solver's stream)"""
    cands = {k: v for k, v in tm.items() if k in SINGLE_KERNEL and v["bytes"] > 0 and v["count"] > 0}
    if not cands:
        return None
    k = max(cands, key=lambda n: cands[n]["ms"])
    v = cands[k]
    gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9
    kname = "%s<%s>" % ("k_helm_" + HELM["mode"] if k == "helm" else SINGLE_KERNEL[k], storage)
    traffic = measured_traffic(kname, workload)
    avg_s = v["ms"] / v["count"] * 1e-3
    out = dict(kernel=kname, bound="hbm", achieved=round(gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
               frac=round(gbs / HBM_PEAK_GBS, 4), traffic=traffic,
               launches=v["count"], avg_us=round(avg_s * 1e6, 2),
               bytes_per_launch=v["bytes"] / v["count"])
    if traffic:
        # the measured HBM bytes over the same launch time, beside what a plain
        # 1 GiB device copy sustains on this part
        out["traffic_GBs"] = round(traffic / avg_s / 1e9, 1)
    out["copy_GBs"] = copy_bandwidth()
    return out


def copy_bandwidth():
    """bytes read + written per second of a plain 1 GiB device copy (20 copies
    after 5 warm-up copies): the HBM rate sustained outside this library"""
    import torch
    x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    t0, t1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0.record()
    for _ in range(20):
        y.copy_(x)
    t1.record()
    torch.cuda.synchronize()
    sec = t0.elapsed_time(t1) / 1e3 / 20
    del x, y
    return round(2 * (1 << 30) / sec / 1e9, 1)


# discrete-adjoint workloads: the 256x96x64 case (FP64, 200 steps, 20
# checkpoints; the gradient-check case) and the 1024x160x192 case (100 steps,
# 10 checkpoints); objective stations x in [0.6, 0.9] Lx, control x in [0.2, 0.4] Lx
ADJ_CASES = {"256x96x64": "c4", "1024x160x192": "c5b"}


def run_adjoint(case="c4", store_all=True):
    """time of the window's forward (J only) and of the full gradient (forward
    sweep + binomial reverse sweep) on the same start state and control, at
    the case's precision: FP64 (256x96x64), or mixed forward with the adjoint
    accumulated in FP64 (1024x160x192)"""
    from turbulentbl_amd import cases, tbl
    c = cases.CASES[case]
    steps, nsnap = c["steps"], c["nsnap"]
    nX, nY, nZ = c["grid"]
    inp = make_inputs(case)
    x = inp["mesh"]["x1"]
    rng = np.random.default_rng(c["seed"])
    vw = np.zeros((nX, nY), order="F")
    sel = (x >= 0.2 * x[-1]) & (x <= 0.4 * x[-1])
    vw[sel, 1:nY - 1] = 1e-3 * inp["Ue"] * rng.standard_normal((int(sel.sum()), nY - 2))
    i0, i1 = int(0.6 * (nX - 1)) + 1, int(0.9 * (nX - 1)) + 1

    prec = c["precision"]

    def fresh(**kw):
        return make_solver(inp, prec, 40, 1e-10, 1e-12 if prec == "fp64" else 1e-4, **kw)
    out = {}
    grads = {}
    # the binomial schedule with nsnap checkpoints, and every state kept in HBM
    # (nsnap = steps: each forward step runs once)
    modes = [("forward", nsnap), ("gradient", nsnap)] + ([("gradient_all", steps)] if store_all else [])
    for mode, ns in modes:
        sv = fresh()
        sv.adjoint_window(2, i0, i1, wall_w=vw, forward_only=True)       # warm-up (library load, first launches)
        sv.close()
        sv = fresh()
        sv.sync()
        t0 = time.perf_counter()
        J, g, nf, na = sv.adjoint_window(steps, i0, i1, wall_w=vw, nsnap=ns, forward_only=(mode == "forward"))
        sv.sync()
        dt = time.perf_counter() - t0
        dev_bytes = sv.device_bytes
        sv.close()
        grads[mode] = g
        out[mode] = dict(seconds=dt, J=J, nfwd=nf, nadj=na, gnorm=None if g is None else float(np.abs(g).max()),
                         device_bytes=dev_bytes)
        log("adjoint %s (%d checkpoints): %.3f s, J %.12e, forward steps %d, adjoint steps %d"
            % (mode, ns, dt, J, nf, na))
    # the library's default step at this precision (z-line Gauss-Seidel, FGMRES + fast preconditioner)
    def default_step(precision):
        sv = make_solver(inp, precision, 40, 1e-10, 1e-12 if precision == "fp64" else 1e-4)
        sv.step(2)
        sv.sync()
        t0 = time.perf_counter()
        sv.step(20)
        sv.sync()
        t = (time.perf_counter() - t0) / 20
        sv.close()
        return t
    t_def = default_step(prec)
    t_def64 = t_def if prec == "fp64" else default_step("fp64")
    f, gr = out["forward"], out["gradient"]
    es = 8 if prec == "fp64" else 4
    N = nX * nY * nZ
    slot_gb = (4 * es + 8) * N / 1e9                                      # Ut, Press (storage type), PCorr (FP64)
    rec = dict(workload="discrete adjoint %s (nX=%d, nY=%d span, nZ=%d wall-normal), %s, %d steps, %d "
                        "checkpoints, objective x in [0.6,0.9]Lx, control x in [0.2,0.4]Lx"
                        % (c["name"], nX, nY, nZ,
                           "FP64" if prec == "fp64" else "mixed-precision forward, FP32 checkpoints, adjoint "
                                                         "accumulated in FP64", steps, nsnap), physics=inp["phys"],
                forward_s=round(f["seconds"], 4), gradient_s=round(gr["seconds"], 4),
                adjoint_over_forward=round(gr["seconds"] / f["seconds"], 3),
                forward_steps_run=gr["nfwd"], adjoint_steps_run=gr["nadj"],
                ms_per_forward_step=round(f["seconds"] / steps * 1e3, 3),
                ms_per_default_step=round(t_def * 1e3, 3),
                adjoint_over_default_forward=round(gr["seconds"] / (t_def * steps), 3),
                ms_per_default_fp64_step=round(t_def64 * 1e3, 3),
                adjoint_over_default_fp64_forward=round(gr["seconds"] / (t_def64 * steps), 3), J=f["J"],
                # HBM: the context, plus the window's buffers: nsnap + 2 state slots
                # and 18 FP64 adjoint fields (21 in a mixed context: the FP64 copy
                # of the linearisation velocity)
                device_gb_context=round(gr["device_bytes"] / 1e9, 2),
                device_gb_window=round((nsnap + 2) * slot_gb + (18 if prec == "fp64" else 21) * N * 8 / 1e9, 2))
    if store_all:
        ga = out["gradient_all"]
        # every state of the window kept in HBM (nsnap = steps); the gradient
        # does not depend on the checkpoint count
        rec.update(gradient_store_all_s=round(ga["seconds"], 4),
                   adjoint_over_forward_store_all=round(ga["seconds"] / f["seconds"], 3),
                   forward_steps_run_store_all=ga["nfwd"],
                   device_gb_window_store_all=round((steps + 2) * slot_gb + 18 * N * 8 / 1e9, 2),
                   store_all_max_abs_diff=float(np.abs(grads["gradient_all"] - grads["gradient"]).max()))
    return rec


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(inp, steps=1, warmup=1):
    """The reference's CPU path on the benchmark workload: the oracle (FP64 C
    restatement with the reference's orderings: per-slab lexicographic
    Gauss-Seidel with halo snapshots, per-slab block SSOR-preconditioned
    GMRES to the reference's errtol 1e-5 rule) on the full grid at the case's
    physics, decomposed into P wall-normal slabs as the reference's MPI ranks
    (P = the OpenMP threads, one slab per thread, threads pinned to cores).
    One warm-up step (first touch, PCorr = 0 start), then `steps` timed steps."""
    os.environ.setdefault("OMP_PROC_BIND", "close")
    os.environ.setdefault("OMP_PLACES", "cores")
    from oracle import orc
    from turbulentbl_amd import cases
    nX, nY, nZ = cases.CASES[inp["case"]]["grid"]
    threads = int(orc.lib().orc_threads())
    P = max(1, min(threads, nZ // 3, 256))
    kw = cases.solver_params(inp["case"], inp["dt"], inp["Re"], inp["Ue"])
    g = orc.Grid(nX, nY, nZ, P=P, **{k: kw[k] for k in ("Lx", "Ly", "Lz", "delta", "dt", "Re", "Ue", "i_rec",
                                                         "planes")})
    s = orc.State(g)
    s.Ut[...] = inp["Ut"]
    rs = inp["rs"]
    s.avg[...] = rs["avg"]; s.Wt[...] = rs["Wt"]; s.newk[...] = rs["newk"]; s.Ume[...] = rs["Ume"]
    s.Ufluc_rec[...] = inp["uf"]
    s.nstep = 1000
    for _ in range(warmup):
        s.step()
    t0 = time.perf_counter()
    ds = [s.step() for _ in range(steps)]
    dt = time.perf_counter() - t0
    return dict(value=round(nX * nY * nZ * steps / dt / 1e6, 4), unit="Mcell-steps/s", cores=threads,
                kind="port", cpu=cpu_model(), slabs=P,
                sample="the %s case at its physics (%dx%dx%d, N = %d), FP64 oracle with the reference's "
                       "orderings and GMRES(50) errtol 1e-5 rule, %d wall-normal slabs on %d pinned threads; "
                       "%d warm-up step, %d timed: %.1f s, GMRES %s it, GS sweeps %s"
                       % (cases.CASES[inp["case"]]["name"], nX, nY, nZ, nX * nY * nZ, P, threads, warmup,
                          steps, dt, [d["gmres"]["iters"] for d in ds], [int(x) for x in ds[-1]["gs_sweeps"]]),
                seconds=round(dt, 2))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    # three warm-up steps: the first pressure solves start from PCorr = 0 and
    # take three refinement cycles; from the third step on the previous
    # step's PCorr is a close enough initial guess for two
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="c3", choices=["c1", "c2", "c3", "c5a"])
    ap.add_argument("--precision", default="both", choices=["both", "mixed", "fp64"])
    ap.add_argument("--kmax", type=int, default=40)
    ap.add_argument("--errtol", type=float, default=1e-10)
    ap.add_argument("--pc-extrap", type=int, default=0, choices=[0, 1],
                    help="pressure-solve initial guess: 0 = previous PCorr, 1 = 2 PCorr(n-1) - PCorr(n-2)")
    ap.add_argument("--helm", default="zebra", choices=sorted(HELM_MODES),
                    help="Gauss-Seidel ordering of the diffusion solve: z-line zebra over span rows (default) "
                         "or z-line red-black over columns")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--out", default=None, help="also write the full record (JSON) here")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the discrete-adjoint cost measurement")
    ap.add_argument("--no-reference-orderings", action="store_true",
                    help="skip the reference's own algorithm (per-slab orderings) on the GPU")
    ap.add_argument("--reference-orderings-only", type=int, default=0, metavar="P",
                    help="only the reference's own algorithm on the GPU, on P slabs")
    ap.add_argument("--adjoint-only", action="store_true", help="only the discrete-adjoint cost measurement")
    ap.add_argument("--adjoint-case", default="256x96x64", choices=sorted(ADJ_CASES),
                    help="adjoint workload for --adjoint-only")
    ap.add_argument("--long-compare", type=int, default=0, metavar="STEPS",
                    help="only compare mixed against FP64 over STEPS steps (checkpoints every 100)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a plain 1 GiB device copy 5 times (HBM counter calibration) and exit")
    a = ap.parse_args()
    HELM["mode"] = a.helm
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: the path holds the whole domain on one MI355X; "
              "multi-GPU slab sharding is design-only (see DESIGN.md)", file=sys.stderr)
        sys.exit(2)
    import torch
    assert torch.cuda.is_available(), "bench.py needs the GPU"
    if a.calibrate_copy:
        x = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        y = torch.empty_like(x)
        for _ in range(5):
            y.copy_(x)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": x.numel() * 4, "copy_bytes_written": y.numel() * 4, "launches": 5}))
        return
    from turbulentbl_amd import tbl
    tbl.load()
    from turbulentbl_amd import cases
    if a.adjoint_only:
        print(json.dumps({"adjoint": run_adjoint(ADJ_CASES[a.adjoint_case])}))
        return
    cfg = cases.CASES[a.config]
    nX, nY, nZ = cfg["grid"]
    inp = make_inputs(a.config)
    if a.reference_orderings_only:
        print(json.dumps({"reference_orderings": run_reference_orderings(inp, a.reference_orderings_only)}))
        return
    if a.long_compare:
        rows = run_long_compare(inp, a.long_compare, min(100, a.long_compare), a.kmax, a.errtol)
        print(json.dumps({"long_compare": {"workload": cfg["name"], "physics": inp["phys"], "steps": a.long_compare,
                                           "max_cf_mean_rel": max(r["cf_mean_rel"] for r in rows),
                                           "max_cf_station_rel": max(r["cf_station_max_rel"] for r in rows),
                                           "max_u_rel": max(r["u_max_rel"] for r in rows),
                                           "checkpoints": rows}}))
        return
    res = {}
    if a.precision in ("both", "fp64"):
        res["fp64"] = run_gpu(inp, "fp64", a.steps, a.warmup, a.kmax, a.errtol, 1e-12,
                              fields=a.precision == "both", pc_extrap=a.pc_extrap)
    if a.precision in ("both", "mixed"):
        res["mixed"] = run_gpu(inp, "mixed", a.steps, a.warmup, a.kmax, a.errtol, 1e-4,
                               fields=a.precision == "both", pc_extrap=a.pc_extrap)
    adj = adj5 = None
    if not a.no_adjoint:
        try:
            adj = run_adjoint("c4")
        except Exception as e:  # the adjoint line must not hide the step result
            adj = dict(error=str(e))
        try:
            adj5 = run_adjoint("c5b", store_all=False)
        except Exception as e:
            adj5 = dict(error=str(e))
    main_key = "mixed" if "mixed" in res else "fp64"
    r = res[main_key]
    cpu = None
    if not a.no_cpu:
        try:
            cpu = cpu_baseline(inp)
        except Exception as e:  # the baseline must not hide the GPU result
            cpu = dict(error=str(e))
    ref = None
    if not a.no_reference_orderings:
        # the CPU baseline's algorithm on the GPU, on the same slab count
        P = cpu["slabs"] if cpu and "slabs" in cpu else min(16, nZ // 3)
        try:
            ref = run_reference_orderings(inp, P)
            if cpu and "value" in cpu:
                ref["over_cpu_same_algorithm"] = round(ref["value"] / cpu["value"], 1)
        except Exception as e:
            ref = dict(error=str(e))
    def step_roofline(x):
        # algorithmic HBM bytes of every timed class over the breakdown steps
        # (each class counts its fields' compulsory reads and writes; the GMRES
        # solution update is not counted), per step, over the timed rate
        B = sum(v["bytes"] for v in x["timers"].values()) / a.steps
        ach = B / (x["ms_per_step"] * 1e-3) / 1e9
        return dict(bound="hbm", bytes_per_step=round(B), ms_per_step=round(x["ms_per_step"], 3),
                    achieved=round(ach, 1), peak=8000.0, unit="GB/s", frac=round(ach / 8000.0, 4),
                    by_class_GB={k: round(v["bytes"] / a.steps / 1e9, 3) for k, v in x["timers"].items()
                                 if v["bytes"]})

    def summary(x):
        dg = x["diags"]
        return dict(value=round(x["rate"], 3), ms_per_step=round(x["ms_per_step"], 3),
                    gmres_iters=[d["gm_iters"] for d in dg], gs_sweeps=[d["gs_sweeps"] for d in dg],
                    true_res=[d["gm_true_res"] for d in dg],
                    # ||pRhs - A PCorr|| / ||pRhs|| after each solve
                    ppe_rel_res=[d["gm_sumav"] for d in dg],
                    # RMS divergence the pressure solve leaves unresolved, dt ||pRhs - A PCorr||_2 / sqrt(N_int)
                    div_unresolved_rms=[x["dt"] * d["gm_true_res"] / np.sqrt(x["nint"]) for d in dg],
                    # the reference's printed 4th-order metric |sum div| / N_int (dnssetbl.f90:851-882)
                    div=[d["div"] for d in dg],
                    roofline=roofline(x["timers_roof"], "float" if x is res.get("mixed") else "double",
                                      cfg["name"]),
                    step_roofline=step_roofline(x),
                    # the same steps' rate with every class timed (event records between launches)
                    ms_per_step_all_classes_timed=round(x["ms_per_step_all_timed"], 3),
                    kernels={k: dict(count=v["count"], ms=round(v["ms"], 3),
                                     GBs=round(v["bytes"] / max(v["ms"], 1e-9) / 1e6, 1))
                             for k, v in x["timers"].items()})
    out = {
        "metric": METRIC,
        "value": round(r["rate"], 3),
        "unit": "Mcell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(r["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if main_key == "mixed" else "f64",
        "data": ("synthetic (seeded Blasius profile + 5% uniform noise)" if cfg["ic"] == "blasius" else
                 "synthetic (seeded log-law + divergence-free Fourier modes, restart recycling state)"),
        "config": {"workload": "TBL fractional step %s (nX=%d stream, nY=%d span, nZ=%d wall-normal), %s; "
                               "Re_theta,in %g (Re_tau %.1f, U_inf/u_tau %.3f), dt %.4g from CFL %g, domain %s, "
                               "tanh stretch %g (first cell y+ %.2f); z-line %s Gauss-Seidel diffusion; "
                               "GMRES(%d) with the fast-diagonalisation preconditioner to true rel. residual %g"
                               % (cfg["name"], nX, nY, nZ, main_key, inp["phys"]["re_theta_in"], inp["Re"], inp["Ue"],
                                  inp["dt"], inp["phys"]["cfl"], inp["phys"]["domain"], inp["phys"]["stretch"],
                                  inp["phys"]["first_cell_yplus"],
                                  "zebra (span rows)" if a.helm == "zebra" else "red-black (columns)", a.kmax, a.errtol),
                   "cells": nX * nY * nZ, "physics": inp["phys"]},
        "roofline": roofline(r["timers_roof"], "float" if main_key == "mixed" else "double", cfg["name"]),
        "step_roofline": step_roofline(r),
        "cpu_baseline": cpu,
        "reference_orderings": ref,
        "adjoint": adj,
        "adjoint_mixed": adj5,
    }
    for k, x in res.items():
        out[k] = summary(x)
    if "mixed" in res and "fp64" in res:
        out["mixed_over_fp64"] = round(res["mixed"]["rate"] / res["fp64"]["rate"], 3)
        tm, tf = res["mixed"]["tau"], res["fp64"]["tau"]
        out["cf_rel_err_mixed_vs_fp64"] = float(abs(tm[1:-1].mean() - tf[1:-1].mean()) / abs(tf[1:-1].mean()))
        um, uf_, pm, pf = res["mixed"]["Ut"], res["fp64"]["Ut"], res["mixed"]["Press"], res["fp64"]["Press"]
        out["u_rel_err_mixed_vs_fp64"] = float(np.abs(um - uf_).max() / np.abs(uf_).max())
        out["p_rel_err_mixed_vs_fp64"] = float(np.abs(pm - pf).max() / np.abs(pf).max())
        dm, df = res["mixed"]["diags"][-1]["div"], res["fp64"]["diags"][-1]["div"]
        out["div_rel_diff_mixed_vs_fp64"] = float(abs(dm - df) / abs(df))
    if a.out:
        with open(a.out, "w") as f:
            json.dump(out, f, indent=1)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
