"""Throughput of the MI355X six-face GAMERA-Cube step.

    python bench.py [--steps K] [--warmup W] [--workload NAME] [--gpus 1]

Workload (default `config1`): 6 faces x 96^2 angular x 192 radial cells,
log radial grid 2.5-30 R_E, uniform solar wind with 1% density/pressure
perturbations, aligned dipole, Earth radial boundary coupled to the
ionosphere (MI) potential solved on two 128 x 256 polar grids (Pedersen
10 S, Hall 5 S) every 10 simulated seconds; `config1_nomi` holds the
potential at zero. The state is resident in HBM before timing starts.

Prints one JSON line: cell-updates/s over K driver steps (CFL dt rule + full
six-face step), the step's fraction of the HBM roofline under the
1216 B/cell-update model, the dominant kernel's roofline from HIP events on
the step's stream next to a measured copy-kernel bandwidth, the CPU rate of
the compiled reference (OpenMP on all host threads; --cpu-kind port: the
oracle) on a two-step sample of the same workload started from the GPU
problem's own arrays, and the rate of the reference's own driver loop
calling the drop-in entry points with the state resident on the device.
"""
import argparse
import ctypes
import importlib
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# Compulsory bytes of one cell-update of the whole step (FP64; 30 doubles of
# state + 122 of static geometry/background).
STEP_BYTES_PER_CELL = 1216
# Algorithmic (compulsory) bytes per unit of each heavy stage: every array
# element the kernel needs, read or written once per launch. Face flux: one
# cell's U*, B*, volume (9 doubles), the face frame, B0, B0.n, Phi* (15),
# fluid + Maxwell flux out (8). Edge EMF: edge frame, solve, B0.t, length
# (14), two transverse faces' area/velocity/Phi* (7), rho* (1), EMF out (1).
# Update: 66 doubles in, U/B/Phi out (11). Predict: Phi, Phi_old, U, U_old,
# face centroids, centroid, recovery constants, volume (38 in); Phi*, U*, B*
# and the primitive, velocity and B.n caches (22 out). The second
# entry is the unit count per launch as a function of (nr, n).
STAGE_MODEL = {
    "face_flux_i": (256, lambda nr, n: 6 * (nr + 1) * n * n),
    "face_flux_j": (256, lambda nr, n: 6 * nr * (n + 1) * n),
    "face_flux_k": (256, lambda nr, n: 6 * nr * n * (n + 1)),
    "edge_emf_i": (184, lambda nr, n: 6 * nr * (n + 1) * (n + 1)),
    "edge_emf_j": (184, lambda nr, n: 6 * (nr + 1) * n * (n + 1)),
    "edge_emf_k": (184, lambda nr, n: 6 * (nr + 1) * (n + 1) * n),
    # the update also folds the new state's CFL dt (3 CFL lengths more)
    "update": (640, lambda nr, n: 6 * nr * n * n),
    "predict": (480, lambda nr, n: 6 * (nr + 9) * (n + 9) * (n + 9)),
}


# The FP64 kernel of each modelled stage (rocprofv3's names, templates kept).
STAGE_KERNEL = {
    "face_flux_i": "k_face_flux_tiled<0, 8, 16>", "face_flux_j": "k_face_flux_tiled<1, 8, 16>",
    "face_flux_k": "k_face_flux", "edge_emf_i": "k_edge_emf_tiled<0>",
    "edge_emf_j": "k_edge_emf_tiled<1>", "edge_emf_k": "k_edge_emf",
    "update": "k_update<false>", "predict": "k_predict<false>",
}


# The reference's production switches (scripts/run_earth.py:18-20): result-
# preserving caches and the parallel planned halo fill, read with getenv by
# the reference library while it steps.
REFERENCE_FLAGS = ("GAMERA_CS_HALO_COMPACT", "GAMERA_CS_HALO_INDEX_CACHE",
                   "GAMERA_CS_HALO_BLOCK_ACCESS", "GAMERA_CS_HALO_PARALLEL",
                   "GAMERA_NO_EMF_FACE_CACHE", "GAMERA_NO_PRIMITIVE_CACHE")


def copy_of(lib, prefix, gpu_problem, cfg, prob):
    """A problem of `lib` (the compiled reference through its harness,
    prefix "refh", allocated without the reference's setup; or the oracle,
    "orc") holding the GPU problem's geometry, background, state and clock,
    without the MI coupling."""
    set_clock = getattr(lib, f"{prefix}_set_clock")
    set_clock.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double,
                          ctypes.c_double, ctypes.c_long]
    ocfg = type(cfg).from_buffer_copy(cfg)
    ocfg.mi_enabled = 0
    if prefix == "refh":
        o = prob.Problem(lib, prefix, ocfg, create="create_shell")
    else:
        o = prob.Problem(lib, prefix, ocfg)
    o.copy_fields(gpu_problem, prob.GEOMETRY_FIELDS + prob.BACKGROUND_FIELDS
                  + prob.STATE_FIELDS)
    for f in range(6):
        for d in range(3):
            o.edge_valid(f, d)[:] = gpu_problem.edge_valid(f, d)
    c = gpu_problem.clock()
    set_clock(o._handle, c["time"], c["dt"], c["previous_dt"], c["steps"])
    return o


def dropin_resident(gpu_problem, cfg, steps, warmup=3):
    """The reference's own driver loop (oracle/_ref harness: the serial Earth
    driver's dt rule and MI cadence) on a copy of the GPU problem, calling
    libgcs.so's gamera_cs_advance_six_with_background_planned,
    gamera_no_local_timestep and gamera_cs_mi_dipole_* in resident mode: the
    state stays on the device, and each step's dt comes from six device
    reductions read back to the host, as the reference's driver reads its
    local dt. Timed over `steps` driver steps after `warmup` (mirror build and
    first upload)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libgamera_ref.so")
    if not os.path.exists(path):
        return None
    prob = importlib.import_module("gamera-cube_amd.problem")
    native = importlib.import_module("gamera-cube_amd.native")
    gcs = native.library()
    ref = ctypes.CDLL(path, mode=ctypes.RTLD_GLOBAL)  # Earth callbacks found by name
    vp = ctypes.c_void_p
    for name in ("refh_grid", "refh_storage", "refh_background", "refh_startup"):
        getattr(ref, name).restype = vp
        getattr(ref, name).argtypes = [vp]
    ref.refh_set_step.argtypes = [vp, vp]
    ref.refh_set_timestep.argtypes = [vp, vp]
    ref.refh_use_mi.argtypes = [vp, vp, vp]
    gcs.gcs_dropin_release.argtypes = [vp, vp]
    gcs.gcs_dropin_sync_to_host.argtypes = [vp, vp]
    gcs.gcs_dropin_host_round_trips.restype = ctypes.c_long
    gcs.gcs_dropin_local_timestep_device_calls.restype = ctypes.c_long
    gcs.gamera_cs_mi_dipole_create.argtypes = [vp, vp, vp, vp, vp, ctypes.c_size_t, vp,
                                               ctypes.POINTER(vp)]
    gcs.gamera_cs_mi_dipole_destroy.argtypes = [vp]

    def addr(name):
        return ctypes.cast(getattr(gcs, name), vp).value

    o = copy_of(ref, "refh", gpu_problem, cfg, prob)
    h = vp()
    grid, storage = ref.refh_grid(o._handle), ref.refh_storage(o._handle)
    try:
        ref.refh_set_step(o._handle, addr("gamera_cs_advance_six_with_background_planned"))
        ref.refh_set_timestep(o._handle, addr("gamera_no_local_timestep"))
        if cfg.mi_enabled:
            active = (ctypes.c_size_t * 3)(cfg.radial_cells, cfg.angular_cells,
                                           cfg.angular_cells)
            if gcs.gamera_cs_mi_dipole_create(grid, storage, ref.refh_background(o._handle),
                                              ref.refh_startup(o._handle), active, cfg.ghost,
                                              ctypes.byref(cfg.mi), ctypes.byref(h)) != 0:
                return {"error": "gamera_cs_mi_dipole_create failed"}
            ref.refh_use_mi(o._handle, h, addr("gamera_cs_mi_dipole_maybe_update"))
        gcs.gcs_dropin_set_resident(1)
        o.advance(warmup)
        gcs.gcs_device_synchronize()
        rounds = gcs.gcs_dropin_host_round_trips()
        served = gcs.gcs_dropin_local_timestep_device_calls()
        t0 = time.perf_counter()
        o.advance(steps)
        gcs.gcs_device_synchronize()
        seconds = time.perf_counter() - t0
        t1 = time.perf_counter()
        if gcs.gcs_dropin_sync_to_host(grid, storage) != 0:
            return {"error": "gcs_dropin_sync_to_host failed"}
        sync_s = time.perf_counter() - t1
        # the default (strict) mode: every call uploads the host state and
        # downloads every array the step writes, over PCIe
        gcs.gcs_dropin_set_resident(0)
        strict_steps = 2
        t2 = time.perf_counter()
        o.advance(strict_steps)
        gcs.gcs_device_synchronize()
        strict_s = time.perf_counter() - t2
        return {"ms_per_step": seconds * 1e3 / steps,
                "strict_ms_per_step": strict_s * 1e3 / strict_steps,
                "strict_cell_updates_per_s": cfg.active_cells * strict_steps / strict_s,
                "cell_updates_per_s": cfg.active_cells * steps / seconds,
                "steps": steps, "warmup": warmup,
                "dt_reductions_on_device": gcs.gcs_dropin_local_timestep_device_calls() - served,
                "host_round_trips": gcs.gcs_dropin_host_round_trips() - rounds,
                "sync_to_host_s": sync_s,
                "mi": ("libgcs.so's gamera_cs_mi_dipole_* (the reference interface has "
                       "no Hall term: Pedersen only)") if cfg.mi_enabled else "off",
                "loop": "the reference's driver loop (harness) calling the drop-in entry "
                        "points: dt from gamera_no_local_timestep per face, MI cadence, "
                        "planned six-face step; state resident on the device"}
    finally:
        gcs.gcs_dropin_set_resident(0)
        ref.refh_use_mi(o._handle, None, None)
        if h.value:
            gcs.gamera_cs_mi_dipole_destroy(h)
        gcs.gcs_dropin_release(grid, storage)
        o.close()


def cpu_baseline(gpu_problem, cfg, steps=2, kind="reference"):
    """The CPU step on the same workload, started from the GPU's arrays:
    kind "reference" is the reference library compiled from its sources
    (oracle/_ref, with its production switches on), "port" the oracle (the
    C restatement). Both run the driver loop without the MI coupling, whose
    solve comes once per ~360 steps at config 1 and would dominate a
    two-step sample; the inner wall is the GPU problem's, copied in."""
    name = {"reference": os.path.join("_ref", "libgamera_ref.so"),
            "port": "liboracle.so"}[kind]
    path = os.path.join(ROOT, "oracle", name)
    if not os.path.exists(path):
        return None
    prob = importlib.import_module("gamera-cube_amd.problem")
    lib = ctypes.CDLL(path)
    prefix = "refh" if kind == "reference" else "orc"
    if kind == "reference":
        for flag in REFERENCE_FLAGS:
            os.environ[flag] = "1"
    o = copy_of(lib, prefix, gpu_problem, cfg, prob)
    o.advance(1)  # first step builds the halo plans: not timed
    t0 = time.perf_counter()
    o.advance(steps)
    seconds = time.perf_counter() - t0
    o.close()
    cores = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    n, nr = cfg.angular_cells, cfg.radial_cells
    what = ("the reference library compiled from its sources (oracle/_ref), OpenMP, "
            "its production cache and halo switches on" if kind == "reference" else
            "the oracle, a C restatement of the reference step, OpenMP")
    return {"value": cfg.active_cells * steps / seconds, "unit": "cell-updates/s",
            "cores": cores, "kind": kind,
            "sample": f"{steps} driver steps (after one untimed) of {what}, on the same "
                      f"6x{n}^2x{nr} problem started from the GPU's arrays (no MI "
                      f"update inside the sample), {seconds:.1f} s"}


def cpu_only(workload):
    """The compiled reference (oracle/_ref, its own setup, production
    switches on) and the oracle (copied from it) on the CPU at `workload`:
    cell-updates/s with all OpenMP threads and with one."""
    prob = importlib.import_module("gamera-cube_amd.problem")
    cfgm = importlib.import_module("gamera-cube_amd.config")
    cfg = cfgm.workload(workload)
    for flag in REFERENCE_FLAGS:
        os.environ[flag] = "1"
    gomp = ctypes.CDLL("libgomp.so.1")
    threads = gomp.omp_get_max_threads()
    ref = ctypes.CDLL(os.path.join(ROOT, "oracle", "_ref", "libgamera_ref.so"))
    orc = ctypes.CDLL(os.path.join(ROOT, "oracle", "liboracle.so"))
    t0 = time.perf_counter()
    r = prob.Problem(ref, "refh", cfg)
    setup_s = time.perf_counter() - t0
    o = prob.Problem(orc, "orc", cfg)
    o.copy_fields(r, prob.GEOMETRY_FIELDS + prob.BACKGROUND_FIELDS + prob.STATE_FIELDS)
    for f in range(6):
        for d in range(3):
            o.edge_valid(f, d)[:] = r.edge_valid(f, d)
    r.advance(1)
    o.advance(1)
    out = {"workload": workload, "active_cells": cfg.active_cells,
           "reference_setup_s": setup_s, "threads_all": threads}
    for label, n, steps in (("all", threads, 20), ("one", 1, 4)):
        gomp.omp_set_num_threads(n)
        for kind, p in (("reference", r), ("port", o)):
            t = time.perf_counter()
            p.advance(steps)
            sec = time.perf_counter() - t
            out[f"{kind}_{label}_threads"] = {
                "cell_updates_per_s": cfg.active_cells * steps / sec, "steps": steps,
                "seconds": sec, "threads": n}
            print(f"bench: {kind} {n} threads {steps} steps {sec:.1f} s", file=sys.stderr,
                  flush=True)
    gomp.omp_set_num_threads(threads)
    print(json.dumps({"cpu_only": out}))
    return 0


def copy_peak(gib=2.0, iters=10):
    """HBM bandwidth of a plain device-to-device copy (read + write bytes per
    second, libgcs.so's copy kernel timed with HIP events) on this GPU, beside
    the 8 TB/s specification."""
    native = importlib.import_module("gamera-cube_amd.native")
    lib = native.library()
    lib.gcs_copy_bandwidth.argtypes = [ctypes.c_size_t, ctypes.c_int,
                                       ctypes.POINTER(ctypes.c_double)]
    out = ctypes.c_double()
    if lib.gcs_copy_bandwidth(int(gib * (1 << 30)) // 8, iters, ctypes.byref(out)) != 0:
        raise RuntimeError(f"copy kernel failed: {native.last_error()}")
    return out.value


def active_state(p, cfg, prob):
    """(rho, p, |v|, |B|) of the active cells, and max |div B| (net total
    face flux over volume, test_cubedsphere_dipole_background.c:808-845,
    computed on the device)."""
    import numpy as np
    ni, nj, nk = cfg.extent
    g, nr, n = cfg.ghost, cfg.radial_cells, cfg.angular_cells
    act = (slice(g, g + nr), slice(g, g + n), slice(g, g + n))
    out = {"rho": [], "p": [], "v": [], "b": []}
    gamma = cfg.startup.gamma
    for f in range(6):
        u = p.field(f, prob.CONSERVED).reshape(ni, nj, nk, 5)[act]
        b = p.field(f, prob.CELL_MAGNETIC).reshape(ni, nj, nk, 3)[act]
        rho = u[..., 0]
        v = u[..., 1:4] / rho[..., None]
        out["rho"].append(rho.ravel())
        out["p"].append(((gamma - 1.0) * (u[..., 4] - 0.5 * rho * (v * v).sum(-1))).ravel())
        out["v"].append(np.linalg.norm(v, axis=-1).ravel())
        out["b"].append(np.linalg.norm(b, axis=-1).ravel())
    return {k: np.concatenate(v) for k, v in out.items()}, p.max_div_b()


def mixed_error(args, cfgm, native, prob):
    """The mixed form against the FP64 step on the same problem after
    `args.mixed_error` driver steps: relative L2 / Linf of rho, p, |v|, |B|
    and max |div B| of both. The FP64 step is the oracle's to 1e-12 (parity
    tests), so it stands for the oracle at this size."""
    import numpy as np
    cfg = cfgm.workload(args.workload)
    runs = {}
    for precision in ("fp64", "mixed"):
        g = native.GpuProblem(cfg)
        g.set_fidelity(False)
        g.set_graph(True)
        g.set_precision(precision == "mixed")
        t0 = time.perf_counter()
        done = 0
        while done < args.mixed_error:
            k = min(100, args.mixed_error - done)
            g.advance(k)
            done += k
            print(f"bench: {precision} {done} steps ({time.perf_counter() - t0:.1f} s)",
                  file=sys.stderr, flush=True)
        state, div = active_state(g, cfg, prob)
        mi = g.mi_report() if cfg.mi_enabled else None
        runs[precision] = (state, div, g.clock()["time"], mi)
        g.close()
    ref, mix = runs["fp64"][0], runs["mixed"][0]
    err = {}
    for k in ("rho", "p", "v", "b"):
        d = mix[k] - ref[k]
        err[k] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ref[k])),
                  "rel_linf": float(np.abs(d).max() / np.abs(ref[k]).max())}
    print(json.dumps({"mixed_error": {"workload": args.workload, "steps": args.mixed_error,
                                      "errors": err,
                                      "max_div_b": {"fp64": runs["fp64"][1],
                                                    "mixed": runs["mixed"][1]},
                                      "time_code": {"fp64": runs["fp64"][2],
                                                    "mixed": runs["mixed"][2]},
                                      "mi_updates": {k: (v[3]["updates"] if v[3] else 0)
                                                     for k, v in runs.items()}}}))
    return 0


def large_run(args, cfgm, native, prob):
    """Config 4: 6 x 192^2 x 384 with MI, Bz switching from +5 to -5 nT at
    about step 500 (the code time of 500 first-step dts), run for
    `args.large_run` hipGraph-launched driver steps in FP64 and in the mixed
    form; cell-updates/s of each and the mixed form's error against FP64."""
    import numpy as np
    probe = native.GpuProblem(cfgm.workload("config4"))
    probe.advance(1)
    t_switch = 500.0 * probe.clock()["dt"]
    probe.close()
    results, states = {}, {}
    for precision in ("fp64", "mixed"):
        cfg = cfgm.bz_switch(cfgm.workload("config4"), t_switch, -5.0)
        t0 = time.perf_counter()
        g = native.GpuProblem(cfg)
        setup_s = time.perf_counter() - t0
        g.set_fidelity(False)
        g.set_graph(True)
        g.set_precision(precision == "mixed")
        total_ms, done = 0.0, 0
        while done < args.large_run:
            k = min(100, args.large_run - done)
            ms, _ = g.time_steps(k, profile=False)
            total_ms += ms
            done += k
            print(f"bench: config4 {precision} {done} steps, {total_ms / done:.2f} ms/step",
                  file=sys.stderr, flush=True)
        mi = g.mi_report()
        states[precision] = active_state(g, cfg, prob)
        results[precision] = {
            "cell_updates_per_s": cfg.active_cells * done / (total_ms * 1e-3),
            "ms_per_step": total_ms / done, "setup_s": setup_s,
            "time_code": g.clock()["time"], "mi_updates": mi["updates"],
            "last_mi_iterations": mi["iterations"], "last_mi_ms": mi["solve_ms"]}
        g.close()
    ref, mix = states["fp64"][0], states["mixed"][0]
    err = {}
    for k in ("rho", "p", "v", "b"):
        d = mix[k] - ref[k]
        err[k] = {"rel_l2": float(np.linalg.norm(d) / np.linalg.norm(ref[k])),
                  "rel_linf": float(np.abs(d).max() / np.abs(ref[k]).max())}
    print(json.dumps({"large_run": {
        "workload": "config4 (6x192^2x384, MI 2x128x256 Hall 5 S, Bz +5 -> -5 nT)",
        "active_cells": cfgm.workload("config4").active_cells, "steps": args.large_run,
        "t_switch_code": t_switch, "runs": results, "mixed_vs_fp64": err,
        "max_div_b": {"fp64": states["fp64"][1], "mixed": states["mixed"][1]},
        "step_launch": "hipGraph"}}))
    return 0


def warm_start_report(g, cfg, steps):
    """Train the warm-start U-Net on the GPU for the problem's MI grid, then
    compare Krylov iterations and milliseconds with and without it: on a
    generated current that changes by 30 % between two updates, and on the
    problem's own current against the potential of its last MI update."""
    import numpy as np
    ws = importlib.import_module("gamera-cube_amd.warm_start")
    nt, np_ = cfg.mi.colatitude_points, cfg.mi.azimuth_points
    # the network runs on the polar grid pooled 4 x 4 (32 x 64 at config 1)
    pool = 4 if nt % 32 == 0 and np_ % 32 == 0 else 1
    net = ws.WarmStartNet(nt // pool, np_ // pool, 8, seed=2)
    t0 = time.perf_counter()
    losses = ws.train(g, net, steps=steps, batch=8, learning_rate=1e-3, seed=9)
    train_s = time.perf_counter() - t0
    print(f"bench: warm start trained {steps} steps in {train_s:.1f} s, relative loss "
          f"{losses[:10].mean():.3e} -> {losses[-50:].mean():.3e}", file=sys.stderr, flush=True)
    ws.attach(g, net)
    out = {"network": "4-level U-Net, 32 base channels, bf16/fp32 MFMA",
           "network_grid": f"2 x {nt // pool} x {np_ // pool} (polar grid pooled {pool}x{pool})",
           "train_steps": steps, "train_batch": 8, "train_s": train_s,
           "relative_loss_first": float(losses[:10].mean()),
           "relative_loss_last": float(losses[-50:].mean())}
    # the problem's own current about 8 simulated seconds after its last MI
    # update (the update interval is 10 s) against that update's potential
    g.advance(300)
    pot_prev = g.mi_field(0).reshape(2, nt, np_)
    g.mi_sample_current()
    cur_now = g.mi_field(1).reshape(2, nt, np_)
    cases = {"mhd_current_vs_last_potential": (cur_now, pot_prev)}
    sp, sh = cfg.mi.pedersen_siemens, cfg.hall_siemens
    cur_a, pot_a = ws.make_pairs(g, seed=21, count=2, pedersen=sp, hall=sh)
    cur_d, _ = ws.make_pairs(g, seed=22, count=2, pedersen=sp, hall=sh)
    cases["generated_30pct_change"] = (
        cur_a + 0.3 * cur_d * (np.abs(cur_a).max() / np.abs(cur_d).max()), pot_a)
    for name, (cur, guess) in cases.items():
        plain, it_p, ms_p = ws.solve_current(g, cur, guess, use_network=False)
        learned, it_l, ms_l = ws.solve_current(g, cur, guess, use_network=True)
        out[name] = {"plain_iterations": it_p, "plain_ms": ms_p,
                     "learned_iterations": it_l, "learned_ms": ms_l,
                     "max_potential_difference_rel": float(
                         np.abs(learned - plain).max() / max(np.abs(plain).max(), 1e-300))}
    ws.attach(g, None)
    net.close()
    return out


def logged_warm_start(runs, cfgm, native, held_out=2, samples=40, every=36, lag=10,
                      train_steps=2000):
    """The learned warm start trained on in-run data: `runs` config-1 MHD runs
    (6 x 96^2 x 192, MI on its 2 x 128 x 256 grids, Pedersen 10 S, Hall 5 S)
    over seeded solar-wind variations (warm_start.WIND_RANGES) log the
    field-aligned current every `every` steps (gcs_problem_mi_sample_current);
    the plain FP64 Krylov solve to 1e-10 gives each current's potential.
    Networks train on the samples and on their changes over `lag` samples
    (about the 10 s MI cadence), then `held_out` further runs compare, for
    each current and the potential `lag` samples earlier: the plain solve
    from that previous potential against the solve from the network's
    correction of it."""
    import numpy as np
    ws = importlib.import_module("gamera-cube_amd.warm_start")
    data, winds, target_iters = [], [], []
    keep = None
    t0 = time.perf_counter()
    for r in range(runs + held_out):
        wind = ws.wind_variation(1000 + r)
        cfg = cfgm.earth_problem(192, 96, perturbation=0.01, seed=1 + r, mi=True,
                                 hall_siemens=5.0, **wind)
        g = native.GpuProblem(cfg)
        g.set_fidelity(False)
        g.set_graph(True)
        cur, times = ws.log_currents(g, samples, every)
        pot, iters = ws.solve_targets(g, cur)
        data.append((cur, pot, times))
        winds.append(wind)
        target_iters += iters
        if keep is not None:
            keep.close()
        keep = g
        print(f"bench: logged run {r + 1}/{runs + held_out} ({wind}), "
              f"{time.perf_counter() - t0:.0f} s", file=sys.stderr, flush=True)
    log_s = time.perf_counter() - t0
    train = [ws.hemisphere_pairs(c, p, lag=l) for c, p, _ in data[:runs] for l in (0, lag)]
    cur_t = np.concatenate([t[0] for t in train])
    pot_t = np.concatenate([t[1] for t in train])
    hemi_t = np.concatenate([t[2] for t in train])
    nt, np_ = keep.config.mi.colatitude_points, keep.config.mi.azimuth_points
    out = {"runs": runs, "held_out_runs": held_out, "samples_per_run": samples,
           "steps_between_samples": every, "lag_samples": lag,
           "lag_seconds": float(np.mean([(t[lag:] - t[:-lag]).mean() for _, _, t in data])
                                * cfgm.TIME_NORM_S),
           "winds": winds, "training_pairs": int(cur_t.shape[0]), "logging_s": log_s,
           "target_solve_iterations_mean": float(np.mean(target_iters)), "networks": {}}
    for label, pool, steps in (("pooled_4x4", 4, train_steps), ("full_grid", 1, train_steps // 2)):
        net = ws.WarmStartNet(nt // pool, np_ // pool, 8, seed=3)
        t1 = time.perf_counter()
        losses = ws.train_on(keep, net, cur_t, pot_t, hemi_t, steps=steps, batch=8,
                             learning_rate=1e-3, seed=5)
        train_s = time.perf_counter() - t1
        ws.attach(keep, net)
        plain_it, plain_ms, learn_it, learn_ms, diff = [], [], [], [], 0.0
        for cur, pot, _ in data[runs:]:
            for k in range(lag, samples):
                p, it_p, ms_p = ws.solve_current(keep, cur[k], pot[k - lag], use_network=False)
                q, it_l, ms_l = ws.solve_current(keep, cur[k], pot[k - lag], use_network=True)
                plain_it.append(sum(it_p) / 2)
                learn_it.append(sum(it_l) / 2)
                plain_ms.append(ms_p)
                learn_ms.append(ms_l)
                diff = max(diff, float(np.abs(q - p).max() / max(np.abs(p).max(), 1e-300)))
        ws.attach(keep, None)
        net.close()
        out["networks"][label] = {
            "grid": f"2 x {nt // pool} x {np_ // pool}", "train_steps": steps,
            "train_s": train_s, "relative_loss_first": float(losses[:10].mean()),
            "relative_loss_last": float(losses[-50:].mean()),
            "held_out_solves": len(plain_it),
            "previous_potential_iterations_mean": float(np.mean(plain_it)),
            "previous_potential_ms_mean": float(np.mean(plain_ms)),
            "learned_iterations_mean": float(np.mean(learn_it)),
            "learned_ms_mean": float(np.mean(learn_ms)),
            "learned_fewer_iterations_fraction": float(np.mean(np.array(learn_it) <
                                                               np.array(plain_it))),
            "max_potential_difference_rel": diff}
        print(f"bench: {label}: iterations {np.mean(plain_it):.1f} -> {np.mean(learn_it):.1f}, "
              f"ms {np.mean(plain_ms):.1f} -> {np.mean(learn_ms):.1f}", file=sys.stderr,
              flush=True)
    keep.close()
    print(json.dumps({"logged_warm_start": out}))
    return 0


def measured_traffic(workload, stage):
    """HBM bytes per launch of the stage's kernel from the committed rocprofv3
    FETCH_SIZE/WRITE_SIZE summary of this workload (calibrated against a copy
    kernel), or None when there is none."""
    path = os.path.join(ROOT, "profiles", f"step_{workload}_hbm_traffic.json")
    kernel = STAGE_KERNEL.get(stage)
    try:
        with open(path) as f:
            entry = json.load(f)["kernels"][kernel]
    except (OSError, KeyError, ValueError):
        return None
    return entry["read_bytes"] + entry["write_bytes"]


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="config1")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-kind", choices=("reference", "port"), default="reference",
                    help="CPU baseline: the compiled reference (oracle/_ref) or the "
                         "oracle's restatement")
    ap.add_argument("--cpu-only", metavar="WORKLOAD",
                    help="only time the compiled reference and the oracle on the CPU, "
                         "on WORKLOAD set up by the reference itself, with all threads "
                         "and with one")
    ap.add_argument("--precision", choices=("fp64", "mixed"), default="fp64",
                    help="fp64: the drop-in's arithmetic; mixed: FP32 reconstruction, "
                         "fluxes and EMF with FP64 update, CT, CFL and ionosphere")
    ap.add_argument("--no-mixed", action="store_true",
                    help="skip the mixed-form timing reported beside the FP64 value")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="N",
                    help="only run a plain copy kernel over N doubles (read N, "
                         "write N) three times: the known-size reference for "
                         "rocprofv3 FETCH_SIZE / WRITE_SIZE")
    ap.add_argument("--mixed-error", type=int, default=0, metavar="STEPS",
                    help="only compare the mixed form with the FP64 step after STEPS "
                         "driver steps of the workload and print the errors")
    ap.add_argument("--warm-start", type=int, default=0, metavar="STEPS",
                    help="after the timed steps, train the MI warm-start network for "
                         "STEPS steps and report its iterations/ms against the plain solve")
    ap.add_argument("--no-dropin", action="store_true",
                    help="skip the drop-in resident leg (the reference's driver loop "
                         "calling libgcs.so's entry points)")
    ap.add_argument("--dropin-steps", type=int, default=20,
                    help="driver steps timed in the drop-in resident leg")
    ap.add_argument("--warm-start-logged", type=int, default=0, metavar="RUNS",
                    help="only train the MI warm start on currents logged from RUNS config-1 "
                         "MHD runs over seeded solar-wind variations and compare it with the "
                         "previous-potential start on two held-out runs")
    ap.add_argument("--large-run", type=int, default=0, metavar="STEPS",
                    help="only run config 4 (FP64 and mixed, Bz switch) for STEPS steps")
    args = ap.parse_args()
    if args.cpu_only:
        return cpu_only(args.cpu_only)
    if args.copy_calibration:
        # libgcs.so's copy kernel (k_copy): one untimed and three timed launches
        native = importlib.import_module("gamera-cube_amd.native")
        lib = native.library()
        lib.gcs_copy_bandwidth.argtypes = [ctypes.c_size_t, ctypes.c_int,
                                           ctypes.POINTER(ctypes.c_double)]
        gbs = ctypes.c_double()
        if lib.gcs_copy_bandwidth(args.copy_calibration, 3, ctypes.byref(gbs)) != 0:
            raise RuntimeError(native.last_error())
        print(json.dumps({"copy_calibration_doubles": args.copy_calibration,
                          "bytes_read": 8 * args.copy_calibration,
                          "bytes_written": 8 * args.copy_calibration}))
        return 0
    if args.gpus != 1:
        print("bench.py runs one GPU: the six-face step is a single-GPU "
              "build; --gpus above 1 is not supported", file=sys.stderr)
        return 2
    cfgm = importlib.import_module("gamera-cube_amd.config")
    native = importlib.import_module("gamera-cube_amd.native")
    if args.mixed_error:
        return mixed_error(args, cfgm, native, importlib.import_module("gamera-cube_amd.problem"))
    if args.large_run:
        return large_run(args, cfgm, native, importlib.import_module("gamera-cube_amd.problem"))
    if args.warm_start_logged:
        return logged_warm_start(args.warm_start_logged, cfgm, native)
    cfg = cfgm.workload(args.workload)
    print(f"bench: workload {args.workload} {cfg.active} x 6, setup ...",
          file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    g = native.GpuProblem(cfg)
    g.set_fidelity(False)
    g.set_precision(args.precision == "mixed")
    print(f"bench: setup {time.perf_counter() - t0:.1f} s, warmup {args.warmup}",
          file=sys.stderr, flush=True)
    # Every driver step is one hipGraph launch; the warmup steps capture the
    # graphs of both time-level parities.
    g.set_graph(True)
    if args.warmup:
        g.advance(args.warmup)
    mi_before = g.mi_report()["updates"] if cfg.mi_enabled else 0
    # Timed run: one hipGraph launch per step, no stage events.
    t_before = g.clock()["time"]
    ms, _ = g.time_steps(args.steps, profile=False)
    sim_per_step = (g.clock()["time"] - t_before) / args.steps * cfgm.TIME_NORM_S
    mi_timed = (g.mi_report()["updates"] - mi_before) if cfg.mi_enabled else 0
    print(f"bench: {args.steps} steps in {ms:.1f} ms", file=sys.stderr, flush=True)
    # Stage breakdown: the same number of steps again, kernel by kernel with
    # HIP events between stages on the step's stream.
    ms_profiled, stages = g.time_steps(args.steps, profile=True)
    print(f"bench: profiled {args.steps} steps in {ms_profiled:.1f} ms",
          file=sys.stderr, flush=True)
    per_step = ms / args.steps
    value = cfg.active_cells * args.steps / (ms * 1e-3)
    nr, n = cfg.radial_cells, cfg.angular_cells
    mixed_info = None
    if args.precision == "fp64" and not args.no_mixed:
        # The mixed form on the same problem after the FP64 runs (two
        # warmup steps recapture the graphs), reported beside `value`.
        g.set_precision(True)
        g.advance(2)
        ms_mixed, _ = g.time_steps(args.steps, profile=False)
        g.set_precision(False)
        mixed_info = {"ms_per_step": ms_mixed / args.steps,
                      "cell_updates_per_s": cfg.active_cells * args.steps / (ms_mixed * 1e-3),
                      "speedup_vs_fp64": ms / ms_mixed,
                      "form": "FP32 reconstruction, face fluxes and EMF with FP32 "
                              "intermediate arrays; FP64 update, CT, CFL, ionosphere"}
        print(f"bench: mixed form {args.steps} steps in {ms_mixed:.1f} ms",
              file=sys.stderr, flush=True)
    if cfg.mi_enabled:
        rep = g.mi_report()
        steps_per_update = cfg.mi.coupling_interval_s / sim_per_step
        mi_info = {"grid": f"2 x {cfg.mi.colatitude_points} x {cfg.mi.azimuth_points}",
                   "pedersen_S": cfg.mi.pedersen_siemens, "hall_S": cfg.hall_siemens,
                   "updates_in_timed_steps": mi_timed,
                   "last_solve_iterations": rep["iterations"],
                   "last_solve_ms": rep["solve_ms"],
                   "tolerance": cfg.mi.relative_tolerance,
                   "steps_between_updates": steps_per_update,
                   # the solve's cost spread over the steps of one coupling interval
                   "cell_updates_per_s_with_mi_amortized": cfg.active_cells / (
                       (ms / args.steps + rep["solve_ms"] / steps_per_update) * 1e-3)}
    else:
        mi_info = "off: ionosphere potential held at zero (no-MI wall)"
    peak_copy = copy_peak()
    print(f"bench: copy kernel {peak_copy:.0f} GB/s", file=sys.stderr, flush=True)
    modeled = {k: v for k, v in stages.items() if k in STAGE_MODEL and v > 0}
    top = max(modeled, key=modeled.get)
    bytes_per_unit, units = STAGE_MODEL[top]
    launch_ms = stages[top] / args.steps
    achieved = bytes_per_unit * units(nr, n) / (launch_ms * 1e-3) / 1e9
    result = {
        "metric": "cell-updates/s on 1 MI355X at 6x96^2x192",
        "value": value,
        "unit": "cell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32 flux / f64 update",
        "data": "synthetic",
        "config": {"workload": args.workload,
                   "cells": f"6x{n}^2x{nr}",
                   "active_cells": cfg.active_cells,
                   "mi": mi_info,
                   "precision": args.precision,
                   "mixed_form": mixed_info,
                   "simulated_s_per_step": sim_per_step,
                   "hbm_fraction_step_model": STEP_BYTES_PER_CELL * value / (HBM_PEAK_GBS * 1e9),
                   "hbm_fraction_step_model_of_copy_peak":
                       STEP_BYTES_PER_CELL * value / (peak_copy * 1e9)},
        "roofline": {"bound": "hbm", "kernel": top,
                     "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS,
                     "traffic": measured_traffic(args.workload, top),
                     "traffic_source": f"profiles/step_{args.workload}_hbm_traffic.json: "
                                       "rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of "
                                       "this workload, calibrated on a copy kernel",
                     "measured_copy_peak": peak_copy,
                     "frac_of_measured_copy_peak": achieved / peak_copy,
                     "launch_ms": launch_ms, "bytes_per_unit": bytes_per_unit,
                     "units_per_launch": units(nr, n)},
        "stage_ms_per_step": {k: round(v / args.steps, 4) for k, v in stages.items() if v > 0},
        "profiled_ms_per_step": ms_profiled / args.steps,
        "step_launch": "hipGraph (one captured step per launch)",
        "cpu_baseline": None,
    }
    if args.warm_start and cfg.mi_enabled:
        result["config"]["warm_start"] = warm_start_report(g, cfg, args.warm_start)
    if not args.no_cpu_baseline:
        print("bench: CPU baseline (compiled reference, same workload) ...",
              file=sys.stderr, flush=True)
        result["cpu_baseline"] = cpu_baseline(g, cfg, kind=args.cpu_kind)
    if not args.no_dropin:
        print("bench: drop-in resident leg (reference driver loop, libgcs.so entry "
              "points) ...", file=sys.stderr, flush=True)
        leg = dropin_resident(g, cfg, args.dropin_steps)
        if leg is not None and "cell_updates_per_s" in leg:
            leg["fraction_of_value"] = leg["cell_updates_per_s"] / value
        result["dropin_resident"] = leg
    g.close()
    print(json.dumps(result))
    return 0


if __name__ == "__main__":
    sys.exit(main())
