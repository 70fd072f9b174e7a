'''Train and evaluate the learned initial guess on seeded water clusters.

python -m gpu4pyscf_amd.guess.train_water --train 1000:3000 --eval 5000:5200 \
    --shape 2x2x2 --basis def2-svp --epochs 50 --workers 4 --out model.npz

Training rows come from converged RHF densities of the seeded clusters,
computed by the GPU SCF; with --workers N the seeds are split over N child
processes (started before this process touches the GPU) that each write
their rows to a scratch file.  Prints one JSON line: training loss, and on
the held-out seeds the SCF cycles/time to conv_tol with the learned guess
against the atomic guess, plus the guess-density RMS error against the
converged density.
'''
import argparse
import json
import os
import subprocess
import sys
import tempfile
import time

import numpy as np


def _seeds(spec):
    a, b = spec.split(':')
    return list(range(int(a), int(b)))


def _log(m):
    print(m, file=sys.stderr, flush=True)


def _init_guess(args):
    if not args.init_model:
        return None
    from gpu4pyscf_amd.guess.learned import LearnedGuess
    return LearnedGuess.load(args.init_model)


def _generate(args, shape):
    '''Training rows, in this process or split over child processes.'''
    seeds = _seeds(args.train)
    if args.workers <= 1:
        from gpu4pyscf_amd.guess import learned as L
        return L.make_dataset(seeds, shape, args.spacing, args.basis, log=_log,
                              label_tol=args.label_tol, init_guess=_init_guess(args))
    tmp = tempfile.mkdtemp(prefix='guess_rows_')
    parts = np.array_split(np.asarray(seeds), args.workers)
    procs = []
    for w, part in enumerate(parts):
        if len(part) == 0:
            continue
        out = os.path.join(tmp, f'rows_{w}.npz')
        cmd = [sys.executable, '-m', 'gpu4pyscf_amd.guess.train_water', '--gen-only', out,
               '--train', f'{part[0]}:{part[-1] + 1}', '--shape', args.shape,
               '--spacing', str(args.spacing), '--basis', args.basis,
               '--label-tol', str(args.label_tol)]
        if args.init_model:
            cmd += ['--init-model', args.init_model]
        procs.append((subprocess.Popen(cmd), out))
    Xs, Ys, Ms = [], [], []
    for p, out in procs:
        if p.wait() != 0:
            raise RuntimeError(f'data worker failed with status {p.returncode}')
        d = np.load(out)
        Xs.append(d['X']); Ys.append(d['Y']); Ms.append(d['M'])
        os.remove(out)
    os.rmdir(tmp)
    return np.concatenate(Xs), np.concatenate(Ys), np.concatenate(Ms)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--train', default='1000:1200')
    ap.add_argument('--eval', default='5000:5020')
    ap.add_argument('--shape', default='2x2x2')
    ap.add_argument('--spacing', type=float, default=3.0)
    ap.add_argument('--basis', default='def2-svp')
    ap.add_argument('--epochs', type=int, default=50)
    ap.add_argument('--lr', type=float, default=1e-3)
    ap.add_argument('--conv-tol', type=float, default=1e-8)
    ap.add_argument('--label-tol', type=float, default=1e-9,
                    help='SCF convergence of the training densities')
    ap.add_argument('--workers', type=int, default=1)
    ap.add_argument('--out', default=None)
    ap.add_argument('--init-model', default=None,
                    help='weights of an earlier network whose guess starts the label SCFs '
                         '(fewer cycles; the converged labels do not depend on it)')
    ap.add_argument('--gen-only', default=None, help=argparse.SUPPRESS)
    args = ap.parse_args()
    shape = tuple(int(x) for x in args.shape.split('x'))
    if args.gen_only:
        from gpu4pyscf_amd.guess import learned as L
        X, Y, M = L.make_dataset(_seeds(args.train), shape, args.spacing, args.basis, log=_log,
                                 label_tol=args.label_tol, init_guess=_init_guess(args))
        np.savez(args.gen_only, X=X, Y=Y, M=M)
        return
    t0 = time.perf_counter()
    data = _generate(args, shape)       # before this process initialises the GPU
    t_data = time.perf_counter() - t0
    import torch
    from gpu4pyscf_amd.guess import learned as L
    from gpu4pyscf_amd.scf.hf import RHF
    from gpu4pyscf_amd.scf.atom_guess import atom_guess
    _log(f'[train_water] {data[0].shape[0]} pair rows from {len(_seeds(args.train))} '
         f'geometries in {t_data:.1f} s')
    t0 = time.perf_counter()
    net, hist = L.train(data, epochs=args.epochs, lr=args.lr, log=_log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    guess = L.LearnedGuess(net)
    if args.out:
        guess.save(args.out)
    res = dict(n_train=len(_seeds(args.train)), rows=int(data[0].shape[0]),
               t_data_s=t_data, t_train_s=t_train, final_mse=hist[-1], first_mse=hist[0],
               epochs=args.epochs, workers=args.workers, train_stats=net.train_stats)
    del data
    rows = []
    for s in _seeds(args.eval):
        mol = L.water_mol(s, shape, args.spacing, args.basis)
        out = {}
        for name, key in (('atom', 'atom'), ('learned', guess)):
            mf = RHF(mol)
            mf.conv_tol = args.conv_tol
            mf.init_guess = key
            torch.cuda.synchronize()
            t = time.perf_counter()
            e = mf.kernel()
            torch.cuda.synchronize()
            out[name] = dict(e=e, cycles=mf.cycles, t=time.perf_counter() - t,
                             converged=mf.converged, jk=mf.jk_builds)
            if name == 'atom':
                dconv = mf.dm.cpu().numpy()
        out['rms_learned'] = float(np.sqrt(np.mean((guess(mol) - dconv)**2)))
        out['rms_atom'] = float(np.sqrt(np.mean((atom_guess(mol) - dconv)**2)))
        out['fallback'] = guess.used_fallback
        rows.append(out)
        _log(f"[eval {s}] cycles atom {out['atom']['cycles']} learned {out['learned']['cycles']}"
             f" dE {out['learned']['e'] - out['atom']['e']:.2e}")
    for k in ('atom', 'learned'):
        res[f'cycles_{k}'] = float(np.mean([r[k]['cycles'] for r in rows]))
        res[f'jk_builds_{k}'] = float(np.mean([r[k]['jk'] for r in rows]))
        res[f'time_{k}_s'] = float(np.mean([r[k]['t'] for r in rows]))
        res[f'converged_{k}'] = all(r[k]['converged'] for r in rows)
    res['n_eval'] = len(rows)
    res['max_abs_dE'] = float(max(abs(r['learned']['e'] - r['atom']['e']) for r in rows))
    res['rms_guess_learned'] = float(np.mean([r['rms_learned'] for r in rows]))
    res['rms_guess_atom'] = float(np.mean([r['rms_atom'] for r in rows]))
    res['fallbacks'] = int(sum(r['fallback'] for r in rows))
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
