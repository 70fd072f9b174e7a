'''Benchmark of the direct-SCF J/K engine on one MI355X.

Workload (default, the headline configuration): 64-water cluster (4x4x4
lattice, 3.1 Angstrom spacing, seed 0), def2-TZVP (2752 spherical AOs).

* J/K leg: one step = one incremental get_jk(dD, hermi=1) with
  dD = 1e-3 (G + G^T)/2, G ~ N(0,1) seeded, through the full drop-in
  boundary (AO transforms, dm_cond, all class launches, hermi
  post-processing).  Inputs are resident in HBM before timing; the timed
  steps run the production path (one native call per build, no per-launch
  events).  value = surviving shell quartets per second.
* SCF leg: direct RHF to conv_tol 1e-8 from the same screening setup, plain
  FP64 with the atomic guess against the learned guess + mixed FP32/FP64
  (and the two separately): wall time until the energy stays within 1e-6 Eh
  of the converged plain-FP64 energy, the energy difference and the cycles.
* Roofline: the FP64 class kernels are bound by FP64 VALU issue (no MFMA);
  the per-kernel figure comes from one untimed build with every class
  launch on one stream, so each launch has the GPU to itself.
* CPU baseline: the oracle (C + OpenMP restatement) on a strided sample of
  the same build.

Prints one JSON line.
'''
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def water_mol(shape, spacing, basis, seed=0):
    from gpu4pyscf_amd.gto import geometry as G
    from gpu4pyscf_amd.gto.mole import M
    return M(atom=G.to_string(G.water_cluster(shape, spacing, seed=seed)), basis=basis)


def cpu_baseline(opt, dms_c, target_s=15.):
    '''Oracle (OpenMP C restatement) quartets/s on an ij-strided sample of
    the same J/K build.'''
    from oracle import oracle as O
    sm = opt.sorted_mol
    cache = {k: (v[0], v[1]) for k, v in opt.bas_pair_cache.items()}
    dms = dms_c.cpu().numpy()
    stride = 4096
    t = 0.
    while True:
        t0 = time.perf_counter()
        _, _, n, _ = O.jk_sorted(sm, dms, 1, cache, opt.direct_scf_tol, ij_stride=stride)
        t = time.perf_counter() - t0
        if t > target_s / 4 or stride == 1:
            break
        stride = max(1, int(stride * max(t, 1e-3) / target_s))
    cores = len(os.sched_getaffinity(0))
    return dict(value=n / t, unit='quartets/s', cores=int(os.environ.get('OMP_NUM_THREADS', cores)),
                kind='port',
                sample=f'every {stride}-th ij shell pair of each class task of the same '
                       f'get_jk ({n} quartets, {t:.1f} s)')


def measured_traffic(config, kernel):
    '''HBM bytes per launch of ``kernel`` on ``config`` from the committed
    rocprofv3 FETCH_SIZE/WRITE_SIZE summaries under profiles/ (None when no
    summary covers that kernel and workload).'''
    tag = config.replace('-', '_')
    pdir = os.path.join(ROOT, 'profiles')
    if not os.path.isdir(pdir):
        return None, None
    for fn in sorted(os.listdir(pdir)):
        if tag not in fn or not fn.endswith('.md'):
            continue
        for line in open(os.path.join(pdir, fn)):
            if line.startswith('traffic_bytes_per_launch '):
                name, val = line[len('traffic_bytes_per_launch '):].rsplit('=', 1)
                if name.strip() == kernel:
                    return float(val), f'profiles/{fn}'
    return None, None


def scf_leg(mol, fp32_tol, guess_model=None, conv_tol=1e-8):
    '''Direct RHF to conv_tol: plain FP64 against mixed FP32/FP64 (and the
    learned guess when a model is given); wall times include everything
    after the screening setup.'''
    import torch
    from gpu4pyscf_amd.scf import jk
    from gpu4pyscf_amd.scf.hf import RHF
    from gpu4pyscf_amd.scf.atom_guess import atom_guess
    res = {}
    opt = jk._VHFOpt(mol).build()      # screening setup shared by every variant
    atom_guess(mol)                    # per-element atomic densities, cached
    variants = [('fp64', None, 'atom'), ('mixed', fp32_tol, 'atom')]
    if guess_model:
        from gpu4pyscf_amd.guess.learned import LearnedGuess
        lg = LearnedGuess.load(guess_model)
        variants += [('learned', None, lg), ('learned_mixed', fp32_tol, lg)]
    for name, tol, guess in variants:
        mf = RHF(mol)
        mf.conv_tol = conv_tol
        mf.fp32_tol = tol
        mf.init_guess = guess
        mf._opt = opt
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        e = mf.kernel()
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        res[name] = dict(e_tot=e, wall_s=dt, cycles=mf.cycles, converged=mf.converged,
                         jk_builds=mf.jk_builds, quartets=mf.quartets,
                         fp32_fraction=mf.quartets_fp32 / max(mf.quartets, 1),
                         history=[(float(t), float(x)) for t, x in zip(mf.history_t, mf.history)])
        log(f'[bench] SCF {name}: E={e:.10f} {dt:.2f} s, {mf.cycles} cycles, '
            f'fp32 {res[name]["fp32_fraction"]:.3f}')
    e0 = res['fp64']['e_tot']
    for name in list(res):
        r = res[name]
        r['abs_dE_vs_fp64'] = abs(r['e_tot'] - e0)
        # first SCF energy (after the initial build or a cycle) from which on
        # every energy stays within 1e-6 Eh of the converged plain-FP64 energy
        hist = r.pop('history')
        t_hit = None
        for n in range(len(hist)):
            if all(abs(x - e0) < 1e-6 for _, x in hist[n:]):
                t_hit = (n, hist[n][0])
                break
        r['cycles_to_1e-6'] = t_hit[0] if t_hit else None
        r['time_to_1e-6_s'] = t_hit[1] if t_hit else None
        log(f"[bench] SCF {name}: within 1e-6 Eh of plain FP64 after {r['cycles_to_1e-6']} "
            f"cycles, {r['time_to_1e-6_s']} s")
    t_ref = res['fp64']['time_to_1e-6_s']
    for name in res:
        t = res[name]['time_to_1e-6_s']
        res[name]['speedup_to_1e-6_vs_fp64'] = (t_ref / t) if (t and t_ref) else None
    res['conv_tol'] = conv_tol
    res['fp32_tol'] = fp32_tol
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=5)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--config', default='water64-tzvp')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-scf', action='store_true')
    ap.add_argument('--fp32-tol', type=float, default=1e-6)
    ap.add_argument('--copy-ref', action='store_true',
                    help='also run a 1 GiB device copy (reference for PMC byte counters)')
    ap.add_argument('--guess-model', default='auto',
                    help="learned-guess weights (.npz) for the learned SCF variants; 'auto': "
                         "the shipped weights for the config's basis, 'none': no learned variants")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f'bench.py: this engine runs on one GPU; --gpus {args.gpus} is not supported',
              file=sys.stderr)
        sys.exit(2)
    import torch
    from gpu4pyscf_amd.scf import jk
    from gpu4pyscf_amd.scf.flops import quartet_flops
    if args.config == 'water32-tzvp':
        shape, spacing, basis = (4, 4, 2), 3.0, 'def2-tzvp'
    elif args.config == 'water8-svp':
        shape, spacing, basis = (2, 2, 2), 3.0, 'def2-svp'
    elif args.config == 'water64-tzvp':
        shape, spacing, basis = (4, 4, 4), 3.1, 'def2-tzvp'
    elif args.config == 'alkane30-tzvp':
        shape, spacing, basis = None, None, 'def2-tzvp'
    else:
        raise SystemExit(f'unknown config {args.config}')
    if shape is None:
        # all-trans C30H62, N(0, 0.02 A) jitter, seed 0
        from gpu4pyscf_amd.gto import geometry as G
        from gpu4pyscf_amd.gto.mole import M
        mol = M(atom=G.to_string(G.alkane(30, seed=0)), basis=basis)
    else:
        mol = water_mol(shape, spacing, basis)
    log(f'[bench] {args.config}: nao={mol.nao} nbas={mol.nbas}')
    dev = torch.device('cuda', 0)
    t0 = time.perf_counter()
    opt = jk._VHFOpt(mol).build()
    torch.cuda.synchronize()
    t_build = time.perf_counter() - t0
    log(f'[bench] screening setup {t_build:.2f} s, tasks={len(opt.tasks())}')
    rng = np.random.default_rng(1)
    g = rng.standard_normal((mol.nao, mol.nao))
    ddm = torch.as_tensor(1e-3 * (g + g.T) / 2, device=dev)

    def step():
        return jk.get_jk(mol, ddm, hermi=1, vhfopt=opt)

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    log('[bench] warmup done')
    # timed region: the production path (one native call per build)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / args.steps
    log(f'[bench] get_jk {dt*1e3:.1f} ms/step')
    # one untimed build with per-launch HIP events and quartet counters on
    # the production stream layout
    opt.profile = True
    step()
    prof = opt.profile_results()
    opt.profile = False
    dms_c = opt.apply_coeff_C_mat_CT(ddm[None])
    # one more build (untimed) with every class launch on one stream: the
    # per-launch durations of a kernel that has the GPU to itself
    nstreams, opt.nstreams, opt.profile = opt.nstreams, 1, True
    step()
    torch.cuda.synchronize()
    prof_serial = opt.profile_results()
    opt.nstreams, opt.profile = nstreams, False

    uniq = opt.sorted_mol.uniq_l_ctr

    def per_kernel(prof):
        '''kernel template -> [ms, model flops, launches, quartets]; total flops'''
        total, by = 0., {}
        for t, (ms, cnt, nl) in prof.items():
            ls = tuple(int(uniq[x, 0]) for x in t)
            ps = tuple(int(uniq[x, 1]) for x in t)
            f = cnt * quartet_flops(ls, ps, 1)
            total += f
            a = by.setdefault(jk.class_kernel_name(ls), [0., 0., 0, 0])
            a[0] += ms; a[1] += f; a[2] += nl; a[3] += cnt
        return by, total

    nq = sum(v[1] for v in prof.values())
    kern_ms = sum(v[0] for v in prof.values())
    by_kernel, total_flops = per_kernel(prof)
    by_serial, _ = per_kernel(prof_serial)
    # the dominant kernel template is the one with the largest summed
    # duration in the one-stream build, where each launch has the GPU alone
    dom_s = max(by_serial, key=lambda k: by_serial[k][0])
    s_ms, s_flops, s_n, s_q = by_serial[dom_s]
    serial_ms = sum(v[0] for v in prof_serial.values())
    peak = 78.6
    tflops = lambda f, ms: f / max(ms, 1e-12) / 1e9
    log(f'[bench] quartets={nq} kernels {kern_ms:.1f} ms summed over {opt.nstreams} streams, '
        f'{tflops(total_flops, dt * 1e3):.2f} TFLOP/s model over the wall time')
    log(f'[bench] one-stream build: kernels {serial_ms:.1f} ms; dominant kernel {dom_s}: '
        f'{s_ms:.1f} ms  {tflops(s_flops, s_ms):.2f} TFLOP/s  avg launch {s_ms / s_n * 1e3:.1f} us')
    for key in sorted(by_serial, key=lambda k: -by_serial[k][0])[:40]:
        v = by_serial[key]
        log(f'    {key}: {v[0]:8.2f} ms  {tflops(v[1], v[0]):.2f} TFLOP/s  '
            f'avg launch {v[0] / v[2] * 1e3:.1f} us  launches {v[2]}  quartets {v[3]}')
    for kind in ('jk_lane_kernel', 'jk_team_kernel', 'jk_generic_kernel'):
        sel = [v for k, v in by_serial.items() if k.startswith(kind)]
        if sel:
            ms, fl = sum(v[0] for v in sel), sum(v[1] for v in sel)
            log(f'    all {kind}: {ms:.1f} ms, {tflops(fl, ms):.2f} TFLOP/s, '
                f'{sum(v[3] for v in sel)} quartets')
    # the same incremental build with FP32 quartets below fp32_tol
    vj64, vk64 = opt.get_jk(dms_c, 1)
    opt.get_jk(dms_c, 1, fp32_tol=args.fp32_tol)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        vj32, vk32 = opt.get_jk(dms_c, 1, fp32_tol=args.fp32_tol, count_quartets=True)
    torch.cuda.synchronize()
    dt32 = (time.perf_counter() - t0) / args.steps
    mixed = dict(fp32_tol=args.fp32_tol, ms_per_build=dt32 * 1e3,
                 fp32_fraction=opt.last_quartets_fp32 / max(opt.last_quartets, 1),
                 max_abs_err_j=float((vj32 - vj64).abs().max()),
                 max_abs_err_k=float((vk32 - vk64).abs().max()))
    log(f"[bench] mixed: {dt32*1e3:.1f} ms/step, fp32 fraction {mixed['fp32_fraction']:.3f}, "
        f"max|dJ| {mixed['max_abs_err_j']:.2e} max|dK| {mixed['max_abs_err_k']:.2e}")
    del vj64, vk64, vj32, vk32
    if args.copy_ref:
        # plain device copy of known size: the reference for PMC byte counters
        nbytes = 1 << 30
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(); dst.copy_(src); e1.record()
        torch.cuda.synchronize()
        log(f'[bench] copy reference: {nbytes} B read + {nbytes} B written, '
            f'{e0.elapsed_time(e1):.3f} ms ({2 * nbytes / e0.elapsed_time(e1) / 1e9:.2f} TB/s)')
        del src, dst
    traffic_s, traffic_src = measured_traffic(args.config, dom_s)
    # bound: FP64 vector (VALU) FMA issue; the class kernels issue no MFMA
    roofline = dict(bound='fp64_valu',
                    kernel=f'{dom_s} (all launches of the dominant kernel template, '
                           'one-stream build)',
                    achieved=round(tflops(s_flops, s_ms), 3),
                    peak=peak, unit='TFLOP/s',
                    frac=round(tflops(s_flops, s_ms) / peak, 4), traffic=traffic_s,
                    traffic_unit='bytes per launch (HBM, rocprofv3 FETCH_SIZE + WRITE_SIZE)',
                    traffic_source=traffic_src,
                    avg_launch_us=round(s_ms / s_n * 1e3, 2),
                    flops_per_launch=s_flops / s_n,
                    quartets_per_launch=s_q / s_n,
                    timing='HIP events around each launch of one untimed build after the '
                           'timed region, every class launch on one stream',
                    all_kernels=dict(achieved_wall=round(tflops(total_flops, dt * 1e3), 3),
                                     frac_wall=round(tflops(total_flops, dt * 1e3) / peak, 4),
                                     one_stream_kernel_ms=round(serial_ms, 3),
                                     achieved_one_stream=round(tflops(total_flops, serial_ms), 3),
                                     note='model FLOPs of all class launches over the timed '
                                          'wall time per build (4 streams), and over the summed '
                                          'kernel time of the one-stream build'))
    out = dict(metric='SCF time to 1e-6 Eh, learned+mixed vs plain FP64 (1 GPU); '
                      'J/K quartets/s',
               value=nq / dt, unit='quartets/s', n_gpus=1, steps=args.steps,
               warmup=args.warmup, ms_per_step=dt * 1e3, higher_is_better=True,
               scaling='weak', vs_baseline=None, dtype='f64', data='synthetic',
               config=dict(workload=f'{args.config}: value = one incremental get_jk(dD, '
                                    f'hermi=1) FP64, nao={mol.nao}, quartets={nq}; scf = direct '
                                    f'RHF to 1e-8 Eh, learned+mixed vs plain FP64',
                           global_batch=1, seq_len=0, parallelism='single GPU'),
               roofline=roofline, setup_s=t_build, mixed=mixed)
    if not args.no_scf:
        model = args.guess_model
        if model == 'auto' and args.config.startswith('water'):
            data = os.path.join(ROOT, 'gpu4pyscf_amd', 'guess', 'data')
            model = os.path.join(data, {'def2-tzvp': 'water16_def2tzvp.npz',
                                        'def2-svp': 'water8_def2svp.npz'}[basis])
        elif model == 'auto':
            model = None            # the shipped weights were trained on water
        elif model == 'none':
            model = None
        scf = out['scf'] = scf_leg(mol, args.fp32_tol, model)
        best = 'learned_mixed' if 'learned_mixed' in scf else 'mixed'
        out['scf_headline'] = dict(
            variant=best,
            fp64_time_to_1e6_s=scf['fp64']['time_to_1e-6_s'],
            time_to_1e6_s=scf[best]['time_to_1e-6_s'],
            speedup_vs_fp64=scf[best]['speedup_to_1e-6_vs_fp64'],
            abs_dE_vs_fp64=scf[best]['abs_dE_vs_fp64'],
            fp32_fraction=scf[best]['fp32_fraction'],
            cycles_fp64=scf['fp64']['cycles'], cycles=scf[best]['cycles'])
        log(f"[bench] headline: {best} reaches 1e-6 Eh in {out['scf_headline']['time_to_1e6_s']} s "
            f"vs {out['scf_headline']['fp64_time_to_1e6_s']} s plain FP64 "
            f"({out['scf_headline']['speedup_vs_fp64']}x), |dE| {out['scf_headline']['abs_dE_vs_fp64']:.2e}")
    if not args.no_cpu:
        log('[bench] CPU baseline (oracle) ...')
        out['cpu_baseline'] = cpu_baseline(opt, dms_c)
        log(f"[bench] CPU baseline {out['cpu_baseline']['value']:.3e} quartets/s")
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()
