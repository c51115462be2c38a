#!/usr/bin/env python3
"""Throughput of the CHAPSim2 hot path on one MI355X.

Workload: periodic plane channel 256x128x256 (x, y, z), domain 4pi x 2 x 4pi/3,
two-sided tanh stretching beta = 1.5, CP6 compact schemes, constant x mass
flux, Re_b = 2800, dt = 5e-4; initial field laminar Poiseuille + 10 % seeded
uniform noise.  One step = one RK3 time step (three sub-steps: momentum RHS,
fractional step, Poisson solve, projection).

    python bench.py [--steps K] [--warmup W] [--form fp64c|mixed|fp64r]

Prints one JSON line.  By default both forms of the step run from the same
seeded field: FP64 (`fp64c`) and mixed precision (`mixed`, the headline:
`value` is its Mpts*RK3-steps/s with all fields resident in HBM), followed
by 16 uneven batched channels (`forms.batch`) and the heated channel with
variable properties (`forms.heated`: property refresh from the table and
from the MLP surrogate, each timed, their fields compared after 200 RK3
steps).  `mixed_vs_fp64` holds the
throughput ratio and the mixed form's per-field error against the FP64 form
after the same steps; each form reports max|div u| and its step time with
the reference loop's per-step diagnostics.  `roofline` gives the achieved
HBM bandwidth of the dominant kernel class of the headline form (line
operators along x) from HIP events recorded on the solver's stream over one
further RK3 step in which every kernel runs alone (the timed steps overlap
independent operators on three streams).  `cpu_baseline` times the CPU
oracle (a plain-C OpenMP restatement of the reference solver) on the same
workload on the host cores.  With --form only that form runs.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Mpts·RK3-steps/s at 256×128×256; mixed vs FP64 at max|div u|≤10^-10"
NC = (256, 128, 256)
PEAK_HBM_GBS = 8000.0


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--form", default=None, choices=["fp64c", "mixed", "fp64r"],
                    help="run one form only: fp64c: FP64 with the exact projection (max|div u| at round-off); "
                         "mixed: FP32 momentum right-hand side, FP64 state, divergence and exact projection; "
                         "fp64r: FP64 with the reference's single 2nd-order projection")
    ap.add_argument("--diag-steps", type=int, default=5,
                    help="steps timed with the per-step diagnostics (both-form run)")
    ap.add_argument("--batch-default", type=int, default=16, metavar="N",
                    help="uneven batched channels timed after the two forms (0: none)")
    ap.add_argument("--profile-class", type=int, default=0,
                    help="line-operator direction whose launches are timed (0 x, 1 y, 2 z)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-heated", action="store_true",
                    help="skip the heated variable-property channel (table and MLP property refresh)")
    ap.add_argument("--nc", type=int, nargs=3, default=list(NC))
    ap.add_argument("--batch", type=int, default=0, metavar="N",
                    help="run N independent uneven channels (64^3 .. 192x128x192, mixed precision) together "
                         "as one captured graph per RK3 step instead")
    ap.add_argument("--ops", action="store_true",
                    help="isolated operator benchmark instead: the 18 compact operators, d2 along each axis and "
                         "the Poisson solve, FP64 and FP32, on band-limited inputs (use --nc 512 256 512)")
    ap.add_argument("--tune-to", default=None, metavar="FILE",
                    help="write the launch shapes the autotuner chose (chs_tune_report) to FILE at the end")
    ap.add_argument("--tune-from", default=None, metavar="FILE",
                    help="take the launch shapes listed in FILE (written by --tune-to) instead of timing them, "
                         "so a run under a profiler uses the shapes of an unprofiled run")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (calibrates HBM byte counters)")
    return ap.parse_args()


def tune_from(path):
    """Preset the autotuner with the "key=choice (...)" lines of a report."""
    from chapsim2_amd import _lib

    L = _lib.load()
    for line in open(path):
        key, sep, rest = line.strip().partition("=")
        if sep and rest:
            _lib.check(L.chs_tune_set(key.encode(), int(rest.split()[0])))


def tune_to(path):
    import ctypes as C

    from chapsim2_amd import _lib

    buf = C.create_string_buffer(1 << 20)
    _lib.load().chs_tune_report(buf, len(buf))
    with open(path, "w") as fh:
        fh.write(buf.value.decode())


def copy_probe(mib):
    """A plain copy of known size: 5 torch copies of `mib` MiB, the
    reference point for the FETCH_SIZE / WRITE_SIZE counters."""
    import torch

    n = mib * 1024 * 1024 // 8
    a = torch.rand(n, dtype=torch.float64, device="cuda")
    b = torch.empty_like(a)
    for _ in range(5):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes": n * 8, "copies": 5}))


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_channel_256x128x256.md")


def measured_traffic(form, direction, nc):
    """HBM bytes per launch of the line operators along `direction`, from the
    committed rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload
    (copy-calibrated); None when the summary does not cover the run."""
    if tuple(nc) != NC or not os.path.exists(TRAFFIC_FILE):
        return None
    for line in open(TRAFFIC_FILE):
        cells = [c.strip() for c in line.strip().strip("|").split("|")]
        if len(cells) == 6 and cells[0] == form and cells[1] == direction:
            try:
                return float(cells[5]) * 1e6
            except ValueError:
                return None
    return None


def cpu_baseline(nc, fields, cfg):
    """The oracle's restatement of the reference algorithm (FP64, single
    2nd-order projection per sub-step) on the host cores, on bounded samples
    of the same workload: with every core of this process's share (at most
    16), (T(1 + 2 steps) - T(1 step)) / 2, which leaves out the first step;
    on one core, one sub-step (a third of an RK3 step)."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle

    pyoracle.build()
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    lx, lz = 4 * 3.141592653589793, 4 * 3.141592653589793 / 3
    mpts = nc[0] * nc[1] * nc[2] / 1e6

    def flow():
        of = pyoracle.Flow(list(nc), lx, -1.0, 1.0, lz, wall_y=True, iacc=4, poisson=2, **cfg)
        of.set_state(*fields)
        return of

    pyoracle.set_threads(cores)
    of = flow()
    t = [time.perf_counter()]
    for it in (1, 2, 3):
        of.step(it)
        t.append(time.perf_counter())
    per_step = (t[3] - t[1]) / 2
    del of
    pyoracle.set_threads(1)
    of = flow()
    of.update_re(1)
    t0 = time.perf_counter()
    of.substep(1)
    one = (time.perf_counter() - t0) * 3
    del of
    pyoracle.set_threads(cores)
    return {"value": mpts / per_step, "unit": "Mpts·RK3-steps/s", "cores": cores, "kind": "port",
            "sample": f"the same {nc[0]}x{nc[1]}x{nc[2]} channel with the reference algorithm (FP64, single "
                      f"2nd-order projection): 3 RK3 steps on {cores} threads, timed as (T(3) - T(1)) / 2 = "
                      f"{per_step:.2f} s per step",
            "one_core": {"value": mpts / one, "cores": 1,
                         "sample": f"one sub-step on 1 thread, {one / 3:.1f} s (x3 per RK3 step)"}}


def band_limited(shape, seed, kfrac=1.0 / 3.0):
    """Seeded band-limited random field of `shape` (x, y, z), returned in
    C order (z, y, x): Fourier modes with |k_i| <= N_i * kfrac in every
    direction, random phases, amplitude |k|^(-11/6) (a |k|^(-5/3) shell
    spectrum), zero mean, scaled to max |f| = 1."""
    import numpy as np
    import scipy.fft as sf

    n1, n2, n3 = shape
    rng = np.random.default_rng(seed)
    kx = np.arange(n1 // 2 + 1, dtype=np.float64)
    ky = np.fft.fftfreq(n2, 1.0 / n2)
    kz = np.fft.fftfreq(n3, 1.0 / n3)
    k2 = kz[:, None, None] ** 2 + ky[None, :, None] ** 2 + kx[None, None, :] ** 2
    keep = ((np.abs(kz)[:, None, None] <= n3 * kfrac) & (np.abs(ky)[None, :, None] <= n2 * kfrac) &
            (kx[None, None, :] <= n1 * kfrac) & (k2 > 0))
    amp = np.where(keep, np.power(np.maximum(k2, 1.0), -11.0 / 12.0), 0.0)
    del k2, keep
    spec = amp * np.exp(2j * np.pi * rng.random(amp.shape))
    del amp
    try:
        workers = max(1, min(16, len(os.sched_getaffinity(0))))
    except AttributeError:
        workers = 1
    f = sf.irfftn(spec, s=(n3, n2, n1), workers=workers)
    del spec
    f /= np.abs(f).max()
    return f


def ops_bench(args):
    """Isolated operator benchmark: every compact operator (CP6, channel
    boundary kinds), the second derivative d2 = 1der_P2C after 1der_C2P
    along each axis, the FFT(x,z) + TDMA(y) Poisson solve of a zero-mean
    right-hand side and the 3-D FFT library's channel solve (poisson_010, on a
    uniform y mesh of the same grid), each in FP64 and FP32 on device-resident seeded
    band-limited inputs.  Per kernel: average time from HIP events on the
    solver's stream, algorithmic bandwidth (arrays read + written; five
    read+write passes for the Poisson solve) and the FP32 result's error
    relative to the FP64 one (max |f32 - f64| / max |f64|)."""
    import ctypes as C

    import numpy as np
    import torch

    from chapsim2_amd import _lib
    from chapsim2_amd.domain import FFT_2DECOMP_3DFFT, Domain

    nc = tuple(args.nc)
    dm = Domain(nc, 4 * np.pi, -1.0, 1.0, 4 * np.pi / 3, istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
    L = dm.lib
    names = ["midp_C2P", "midp_P2C", "1der_C2C", "1der_P2P", "1der_C2P", "1der_P2C"]
    inp = {0: "c", 1: "p", 2: "c", 3: "p", 4: "c", 5: "p"}
    out = {0: "p", 1: "c", 2: "c", 3: "p", 4: "p", 5: "c"}
    t0 = time.perf_counter()
    # one field on the y-node extent; cell-centred inputs take its first ny planes
    base = torch.from_numpy(band_limited((nc[0], nc[1] + 1, nc[2]), seed=20261202)).to("cuda")
    gen_s = time.perf_counter() - t0
    vp = C.c_void_p

    def ext(d, loc):
        e = list(nc)
        if d == 1 and loc == "p":
            e[1] = nc[1] + 1
        return e

    def field(e):
        return base[:, : e[1], :].contiguous()

    def bc(d, op):
        return ([5, 5] if inp[op] == "c" else [4, 4]) if d == 1 else [1, 1]

    def op_call(d, op, fi, fo, ei, eo, fbc):
        if fi.dtype == torch.float64:
            return L.chs_op3d(dm.h, d, op, 4, _lib.ivec(bc(d, op)), vp(fi.data_ptr()), _lib.ivec(ei),
                              vp(fo.data_ptr()), _lib.ivec(eo), vp(fbc.data_ptr()) if fbc is not None else None,
                              _lib.MEM_DEVICE)
        return L.chs_op3d_f32(dm.h, d, op, 4, _lib.ivec(bc(d, op)), vp(fi.data_ptr()), _lib.ivec(ei),
                              vp(fo.data_ptr()), _lib.ivec(eo), vp(fbc.data_ptr()) if fbc is not None else None)

    def timed(cls, fn):
        for _ in range(3):
            fn()
        _lib.check(L.chs_profile_begin(dm.h, cls))
        for _ in range(args.steps):
            fn()
        ms, nl, byt = C.c_double(), C.c_longlong(), C.c_double()
        _lib.check(L.chs_profile_end(dm.h, C.byref(ms), C.byref(nl), C.byref(byt)))
        t = ms.value / args.steps  # per application (d2: both launches)
        return round(t * 1e3, 2), round(byt.value / args.steps / (t * 1e-3) / 1e9, 1)

    def rel(a, b):
        return float((a.double() - b).abs().max() / b.abs().max())

    def fbc_for(d, ei, dtype):
        if d != 1:
            return None
        fshape = list(ei)
        fshape[1] = 4
        return torch.zeros(fshape[2], fshape[1], fshape[0], dtype=dtype, device="cuda")

    res = []
    for d in range(3):
        for op in range(6):
            ei, eo = ext(d, inp[op]), ext(d, out[op])
            row = {"op": f"Get_{'xyz'[d]}_{names[op]}_3D"}
            outs = {}
            for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
                fi = field(ei).to(dt)
                fo = torch.zeros(eo[2], eo[1], eo[0], dtype=dt, device="cuda")
                fbc = fbc_for(d, ei, dt)
                torch.cuda.synchronize()
                us, gbs = timed(d, lambda: _lib.check(op_call(d, op, fi, fo, ei, eo, fbc)))
                row[f"{tag}_us"], row[f"{tag}_GBps"] = us, gbs
                outs[tag] = fo
            row["f32_err_vs_f64"] = rel(outs["f32"], outs["f64"])
            res.append(row)
            del outs
    # d2 = 1der_P2C(1der_C2P(f)) along each axis (cell-centred in and out)
    d2 = []
    for d in range(3):
        ec, ep = ext(d, "c"), ext(d, "p")
        row = {"op": f"d2_{'xyz'[d]}"}
        outs = {}
        for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            fi = field(ec).to(dt)
            mid = torch.zeros(ep[2], ep[1], ep[0], dtype=dt, device="cuda")
            fo = torch.zeros(ec[2], ec[1], ec[0], dtype=dt, device="cuda")
            fb1, fb2 = fbc_for(d, ec, dt), fbc_for(d, ep, dt)
            torch.cuda.synchronize()

            def both():
                _lib.check(op_call(d, 4, fi, mid, ec, ep, fb1))
                _lib.check(op_call(d, 5, mid, fo, ep, ec, fb2))

            us, gbs = timed(d, both)
            row[f"{tag}_us"], row[f"{tag}_GBps"] = us, gbs
            outs[tag] = fo
        row["f32_err_vs_f64"] = rel(outs["f32"], outs["f64"])
        d2.append(row)
        del outs
    # Poisson solve of a zero-mean cell-centred right-hand side
    rhs = field(ext(0, "c"))
    rhs -= rhs.mean()

    def time_poisson(dmp, name):
        row = {"op": name}
        outs = {}
        for dt, tag in ((torch.float64, "f64"), (torch.float32, "f32")):
            src = rhs.to(dt)
            work = src.clone()
            fn = (lambda: _lib.check(L.chs_poisson(dmp.h, vp(work.data_ptr()), _lib.MEM_DEVICE))) if tag == "f64" \
                else (lambda: _lib.check(L.chs_poisson_f32(dmp.h, vp(work.data_ptr()))))
            for _ in range(2):
                work.copy_(src)
                torch.cuda.synchronize()
                fn()
            tot_ms, tot_b = 0.0, 0.0
            for _ in range(args.steps):
                work.copy_(src)
                torch.cuda.synchronize()
                _lib.check(L.chs_profile_begin(dmp.h, 3))
                fn()
                ms, nl, byt = C.c_double(), C.c_longlong(), C.c_double()
                _lib.check(L.chs_profile_end(dmp.h, C.byref(ms), C.byref(nl), C.byref(byt)))
                tot_ms += ms.value
                tot_b += byt.value
            t = tot_ms / args.steps
            row[f"{tag}_us"] = round(t * 1e3, 2)
            row[f"{tag}_GBps"] = round(tot_b / args.steps / (t * 1e-3) / 1e9, 1)
            outs[tag] = work
        row["f32_err_vs_f64"] = rel(outs["f32"], outs["f64"])
        return row

    pois = time_poisson(dm, "solve_fft_poisson")
    dm.close()
    # the 3-D FFT library's channel path (poisson_010) on a uniform y mesh of
    # the same grid (the tanh mesh takes FISHPACK, as in the reference)
    pois010 = None
    if nc[1] % 2 == 0:
        dmu = Domain(nc, 4 * np.pi, -1.0, 1.0, 4 * np.pi / 3, istret=0, dt=5e-4, ren=2800.0, idriven=1,
                     poisson=FFT_2DECOMP_3DFFT)
        pois010 = time_poisson(dmu, "poisson_010 (uniform y)")
        dmu.close()
    # launch-shape choices the autotuner made for these kernels (key=choice
    # and the time of each candidate)
    buf = C.create_string_buffer(1 << 16)
    L.chs_tune_report(buf, len(buf))
    tuned = buf.value.decode().splitlines()
    print(json.dumps({"ops_bench": res, "d2": d2, "poisson": pois, "poisson_010": pois010, "grid": list(nc),
                      "inputs": "seeded band-limited random fields (modes |k_i| <= N_i/3, |k|^-5/3 spectrum, "
                                f"seed 20261202, generated in {gen_s:.1f} s); zero-mean Poisson right-hand side",
                      "applications_each": args.steps, "autotuned": tuned}))


def batch_specs(n, seed=20261202):
    """n channel instances drawn from a seed: grids between 64x64x64 and
    192x128x192 (x and z with FFT-friendly sizes, even y), Re_tau in {180, 395}
    (bulk Re 2800 / 6900), tanh stretching beta in [1, 2]."""
    import numpy as np

    from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT

    rng = np.random.default_rng(seed)
    xz = [64, 72, 80, 96, 100, 120, 128, 144, 160, 180, 192]
    ys = list(range(64, 129, 8))
    specs = []
    for _ in range(n):
        beta = float(rng.uniform(1.0, 2.0))
        specs.append(dict(nc=(int(rng.choice(xz)), int(rng.choice(ys)), int(rng.choice(xz))), lxx=4 * np.pi,
                          lyb=-1.0, lyt=1.0, lzz=4 * np.pi / 3, istret=2, rstret=beta / 20.0, dt=5e-4,
                          ren=float(rng.choice([2800.0, 6900.0])), idriven=1, precision=PREC_MIXED,
                          projection=PROJ_EXACT))
    return specs


def batch_run(n, args):
    """n uneven channels (batch_specs) advanced together: one eager RK3 step
    (plans and launch shapes), W warm-up and K timed graph-captured steps;
    every instance's max|div u| afterwards."""
    import torch

    from chapsim2_amd import _lib
    from chapsim2_amd.domain import Batch
    from chapsim2_amd.initial import channel_fields

    specs = batch_specs(n)
    b = Batch(specs)
    for i, (d, s) in enumerate(zip(b.domains, specs)):
        d.set_state(*channel_fields(s["nc"], s["lxx"], -1.0, 1.0, s["lzz"], 2, s["rstret"], noise=0.1, seed=100 + i))
        d.update_re(0)
        d.substep(0)
    b.run(0, max(args.warmup, 2), graph=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    b.run(max(args.warmup, 2), args.steps, graph=True)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    pts = sum(s["nc"][0] * s["nc"][1] * s["nc"][2] for s in specs)
    div = max(float(d.mass_check().max()) for d in b.domains)
    tuned = int(_lib.load().chs_tune_report(None, 0))
    b.close()
    return {"value": pts / 1e6 * args.steps / sec, "ms_per_step": sec * 1e3 / args.steps, "max_div_u": div,
            "what": f"{len(specs)} periodic channels, grids " +
                    ", ".join("x".join(str(v) for v in s["nc"]) for s in specs) +
                    "; Re_b 2800/6900, tanh beta in [1,2], CP6, mixed precision, exact projection, "
                    "one captured graph per RK3 step",
            "total_points": pts, "autotuned_shapes": tuned}


def batch_bench(args):
    r = batch_run(args.batch, args)
    out = {
        "metric": "Mpts·RK3-steps/s, batched channels",
        "value": r["value"], "unit": "Mpts·RK3-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": max(args.warmup, 2), "ms_per_step": r["ms_per_step"], "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": r["what"], "total_points": r["total_points"], "max_div_u": r["max_div_u"],
                   "autotuned_shapes": r["autotuned_shapes"]},
    }
    print(json.dumps(out))


FORMS = {
    "fp64c": "FP64 throughout, exact projection (inverse of the compact D*G per sub-step)",
    "mixed": "FP32 momentum right-hand side (44 line operators, products, fractional step); FP64 state, "
             "bulk-flux sums, divergence, exact projection and correction",
    "fp64r": "FP64 with the reference's single FFT(x,z)+TDMA(y) 2nd-order projection per sub-step",
}


def run_form(form, nc, fields, cfg, args, diag_steps):
    """One form of the channel step: the initial Euler step, W warm-up steps,
    K timed RK3 steps bracketed by device synchronisations, one further step
    with every kernel issued alone and the launches of one line-operator
    class timed with HIP events, then diag_steps steps each followed by the
    reference loop's per-step diagnostics (mass-conservation windows, velocity
    range, CFL, bulk monitor; src/chapsim.f90:319-335).  Returns the result
    record and the final state on the host."""
    import ctypes as C

    import numpy as np
    import torch

    from chapsim2_amd import _lib
    from chapsim2_amd.domain import PREC_FP64, PREC_MIXED, PROJ_EXACT, PROJ_REFERENCE, Domain

    lx, lz = 4 * np.pi, 4 * np.pi / 3
    proj = PROJ_REFERENCE if form == "fp64r" else PROJ_EXACT
    prec = PREC_MIXED if form == "mixed" else PREC_FP64
    dm = Domain(nc, lx, -1.0, 1.0, lz, projection=proj, precision=prec, **cfg)
    dm.set_state(*fields)
    dm.update_re(0)
    dm.substep(0)  # initial Euler step + projection (src/chapsim.f90:135)
    it = 0
    if args.warmup > 0:
        dm.run(it, args.warmup)
        it += args.warmup
    torch.cuda.synchronize()
    L = dm.lib
    t0 = time.perf_counter()
    dm.run(it, args.steps)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    it += args.steps
    # launch timing of one operator class over one further RK3 step: while
    # profiled, the solver issues its kernels one at a time (in the timed
    # steps independent operators overlap on three streams, which stretches
    # each launch's own duration)
    ms, nl, byt = C.c_double(), C.c_longlong(), C.c_double()
    _lib.check(L.chs_profile_begin(dm.h, args.profile_class))
    dm.run(it, 1)
    _lib.check(L.chs_profile_end(dm.h, C.byref(ms), C.byref(nl), C.byref(byt)))
    it += 1
    # steps with the per-step diagnostics of the reference loop
    diag_ms = None
    if diag_steps > 0:
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        for _ in range(diag_steps):
            dm.run(it, 1)
            it += 1
            dm.cfl()
            dm.mass_check()
            dm.velocity_range()
            dm.bulk()
        torch.cuda.synchronize()
        diag_ms = (time.perf_counter() - t2) * 1e3 / diag_steps
    sec = t1 - t0
    mpts = nc[0] * nc[1] * nc[2] / 1e6
    div = float(dm.mass_check().max())
    achieved = (byt.value / max(nl.value, 1)) / (ms.value / max(nl.value, 1) * 1e-3) / 1e9 if nl.value else None
    res = {
        "value": mpts * args.steps / sec,
        "ms_per_step": sec * 1e3 / args.steps,
        "ms_per_step_with_diagnostics": diag_ms,
        "max_div_u": div,
        "steps_total": it,
        "what": FORMS[form],
        "roofline": {
            "bound": "hbm",
            "kernel": f"compact line operators along {'xyz'[args.profile_class]} "
                      f"({'k_line_xu' if args.profile_class == 0 else 'k_line_yz'})",
            "achieved": achieved,
            "peak": PEAK_HBM_GBS,
            "unit": "GB/s",
            "frac": (achieved / PEAK_HBM_GBS) if achieved else None,
            "traffic": measured_traffic(form, "xyz"[args.profile_class], nc),
            "launches": nl.value,
            "avg_launch_ms": ms.value / max(nl.value, 1),
            "measured_on": "one RK3 step after the timed region with every kernel issued alone "
                           "(HIP events on the solver's stream)",
        },
    }
    state = dm.get_state()
    dm.close()
    return res, state


HEATED_STEPS = 200


def heated_run(nc, fields, cfg, args):
    """The heated channel with variable properties (mixed precision, exact
    projection): isothermal walls at h = 1.2, bulk enthalpy 0.8, the
    4096-row table of chapsim2_amd.initial.heated_channel_table on its
    monotone rho*h window, reference Prandtl number 1.  The surrogate
    (1 -> 64 -> 64 -> 5, BF16 weights) is trained on the GPU on 10^6 samples of
    the table's interpolation.  Two solvers start from the same state, one
    refreshing the properties from the table and one from the surrogate;
    each times W warm-up and K steps, then both run on to 200 RK3 steps and
    their fields are compared.  Also times one property refresh of every
    cell in both modes."""
    import numpy as np
    import torch

    from chapsim2_amd.domain import PREC_MIXED, PROJ_EXACT, Domain
    from chapsim2_amd.initial import HEATED_H_RANGE, heated_channel_table
    from chapsim2_amd.mlp import Surrogate
    from chapsim2_amd.props import PROPS_MLP, PROPS_TABLE, PropertyTable

    lx, lz = 4 * np.pi, 4 * np.pi / 3
    tab = heated_channel_table(4096, *HEATED_H_RANGE)
    pt = PropertyTable(tab, -1.0)
    t0 = time.perf_counter()
    sur = Surrogate(pt, nsamples=1_000_000, seed=7)
    loss = sur.train(6000, batch=4096, lr=2e-3)
    torch.cuda.synchronize()
    train_s = time.perf_counter() - t0
    pt.set_surrogate(sur)
    hgrid = np.linspace(0.8, 1.2, 4001)
    ref = pt.from_h(hgrid)
    fit = sur(hgrid)
    fit_err = {k: float(np.abs(fit[k] - ref[k]).max() / np.abs(ref[k]).max()) for k in ("t", "d", "m", "k")}
    tw = float(np.interp(1.2, tab[:, 1], tab[:, 2]))
    ti = float(np.interp(0.8, tab[:, 1], tab[:, 2]))
    mpts = nc[0] * nc[1] * nc[2] / 1e6
    res, doms = {}, {}
    for name, mode in (("table", PROPS_TABLE), ("mlp", PROPS_MLP)):
        dm = Domain(nc, lx, -1.0, 1.0, lz, projection=PROJ_EXACT, precision=PREC_MIXED, **cfg)
        dm.set_state(*fields)
        dm.thermo_enable(pt, (tw, tw), 1.0, ti, mode)
        dm.update_re(0)
        dm.substep(0)
        dm.run(0, args.warmup)
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        dm.run(args.warmup, args.steps)
        torch.cuda.synchronize()
        sec = time.perf_counter() - t1
        done = args.warmup + args.steps
        if done < HEATED_STEPS:
            dm.run(done, HEATED_STEPS - done)
        res[name] = {"value": mpts * args.steps / sec, "ms_per_step": sec * 1e3 / args.steps,
                     "max_div_g": float(dm.mass_check().max()), "steps_total": max(done, HEATED_STEPS)}
        doms[name] = dm
    a, b = doms["mlp"], doms["table"]
    ta, tb = a.thermo_get(), b.thermo_get()
    qa, qb = a.get_state(), b.get_state()
    diff = {k: float(np.abs(ta[k] - tb[k]).max() / np.abs(tb[k]).max()) for k in ("t", "d", "m", "k", "h")}
    diff.update({k: float(np.abs(x - y).max() / np.abs(y).max()) for k, x, y in zip(("qx", "qy", "qz"), qa, qb)})
    hrange = (float(tb["h"].min()), float(tb["h"].max()))
    for dm in doms.values():
        dm.close()
    # one property refresh of every cell, device arrays
    n = nc[0] * nc[1] * nc[2]
    rh = torch.from_numpy(np.interp(np.random.default_rng(1).uniform(0.8, 1.2, n), tab[:, 1],
                                    tab[:, 1] * tab[:, 3])).cuda()
    outs = [torch.empty_like(rh) for _ in range(5)]
    refresh_ms = {}
    for name, mode in (("table", PROPS_TABLE), ("mlp", PROPS_MLP)):
        import ctypes as C

        def call():
            return pt.lib.chs_props_from_rhoh(pt.h, C.c_void_p(rh.data_ptr()), n,
                                              *[C.c_void_p(o.data_ptr()) for o in outs], None, None, mode, 1)
        call()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        for _ in range(5):
            call()
        torch.cuda.synchronize()
        refresh_ms[name] = (time.perf_counter() - t1) * 1e3 / 5
    return {
        "table": res["table"], "mlp": res["mlp"],
        "mlp_vs_table_after_steps": HEATED_STEPS, "mlp_vs_table_rel_diff": diff,
        "h_range_after": hrange,
        "surrogate": {"train_s": train_s, "adam_steps": 6000, "batch": 4096, "samples": 1_000_000,
                      "last_loss": loss, "fit_rel_err_h_0.8_1.2": fit_err},
        "refresh_ms_all_cells": refresh_ms,
        "what": f"heated channel {nc[0]}x{nc[1]}x{nc[2]}, isothermal walls h = 1.2, bulk h = 0.8, "
                "4096-row table over the monotone rho*h window h in [0.65, 1.225], Pr 1, mixed precision, "
                "exact projection; table vs MLP property refresh",
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs the whole domain on one GPU; "
                                   "multi-GPU runs are not part of it (--gpus 1)"}))
        sys.exit(2)
    if args.copy_probe:
        copy_probe(args.copy_probe)
        return
    if args.tune_from:
        tune_from(args.tune_from)
    if args.ops:
        ops_bench(args)
        return
    if args.batch:
        batch_bench(args)
        return
    import numpy as np

    from chapsim2_amd.initial import channel_fields

    nc = tuple(args.nc)
    lx, lz = 4 * np.pi, 4 * np.pi / 3
    cfg = dict(istret=2, rstret=0.075, dt=5e-4, ren=2800.0, idriven=1)
    fields = channel_fields(nc, lx, -1.0, 1.0, lz, 2, 0.075, noise=0.1, seed=20261202)
    workload = (f"periodic channel {nc[0]}x{nc[1]}x{nc[2]}, 4pi x 2 x 4pi/3, tanh beta 1.5, CP6, "
                f"Re_b 2800 constant flux, dt 5e-4, laminar + 10% noise (seed 20261202)")
    forms = [args.form] if args.form else ["fp64c", "mixed"]
    diag = 0 if args.form else args.diag_steps
    results, states = {}, {}
    for form in forms:
        results[form], states[form] = run_form(form, nc, fields, cfg, args, diag)
    head = forms[-1]  # the headline form: mixed when both run
    r = results[head]
    out = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "Mpts·RK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if head == "mixed" else "f64",
        "data": "synthetic",
        "config": {"workload": workload, "form": head, "max_div_u": r["max_div_u"],
                   "div_bar_1e-10_met": bool(r["max_div_u"] <= 1e-10)},
        "roofline": r["roofline"],
    }
    if len(forms) > 1:
        a, b = states["mixed"], states["fp64c"]
        errs = {}
        for name, x, y in zip(("qx", "qy", "qz", "pr"), a, b):
            if name == "pr":  # the pressure is defined up to a constant
                x, y = x - x.mean(), y - y.mean()
            errs[name] = float(np.abs(x - y).max() / np.abs(y).max())
        out["mixed_vs_fp64"] = {
            "ratio": results["mixed"]["value"] / results["fp64c"]["value"],
            "rel_err_per_field": errs,
            "after_steps": results["mixed"]["steps_total"],
            "err_definition": "max |q_mixed - q_fp64| / max |q_fp64| per field after the same steps from the "
                              "same initial field (pressure mean removed)",
        }
        if args.batch_default > 0:
            results["batch"] = batch_run(args.batch_default, args)
            out["mixed_vs_fp64"]["batch_ratio"] = results["batch"]["value"] / results["fp64c"]["value"]
        if not args.no_heated:
            results["heated"] = heated_run(nc, fields, cfg, args)
    out["forms"] = {k: {kk: vv for kk, vv in v.items() if kk != "roofline" or k != head} for k, v in
                    results.items()}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(nc, fields, cfg)
    if args.tune_to:
        tune_to(args.tune_to)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
