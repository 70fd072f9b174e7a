'''Train and evaluate the learned initial guess on seeded water clusters.

python -m gpu4pyscf_amd.guess.train_water --train 0:200 --eval 5000:5020 \
    --shape 2x2x2 --basis def2-svp --epochs 50 --out gpurun_out/mlp_water8_svp.npz

Prints one JSON line: training loss, and on the held-out seeds the SCF
cycles/time to conv_tol with the learned guess against the atomic guess,
plus the guess-density RMS error against the converged density.
'''
import argparse
import json
import sys
import time

import numpy as np


def _seeds(spec):
    a, b = spec.split(':')
    return list(range(int(a), int(b)))


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
    ap.add_argument('--out', default=None)
    args = ap.parse_args()
    import torch
    from gpu4pyscf_amd.guess import learned as L
    from gpu4pyscf_amd.scf.hf import RHF
    log = lambda m: print(m, file=sys.stderr, flush=True)
    shape = tuple(int(x) for x in args.shape.split('x'))
    t0 = time.perf_counter()
    data = L.make_dataset(_seeds(args.train), shape, args.spacing, args.basis, log=log)
    t_data = time.perf_counter() - t0
    log(f'[train_water] {data[0].shape[0]} pair rows from {len(_seeds(args.train))} '
        f'geometries in {t_data:.1f} s')
    t0 = time.perf_counter()
    net, hist = L.train(data, epochs=args.epochs, lr=args.lr, log=log)
    torch.cuda.synchronize()
    t_train = time.perf_counter() - t0
    guess = L.LearnedGuess(net)
    if args.out:
        guess.save(args.out)
    res = dict(n_train=len(_seeds(args.train)), rows=int(data[0].shape[0]),
               t_data_s=t_data, t_train_s=t_train, final_mse=hist[-1], first_mse=hist[0])
    rows = []
    for s in _seeds(args.eval):
        mol = L.water_mol(s, shape, args.spacing, args.basis)
        out = {}
        for name, key in (('atom', 'atom'), ('learned', guess)):
            mf = RHF(mol)
            mf.conv_tol = args.conv_tol
            mf.init_guess = key
            t = time.perf_counter()
            e = mf.kernel()
            out[name] = dict(e=e, cycles=mf.cycles, t=time.perf_counter() - t,
                             converged=mf.converged, jk=mf.jk_builds)
            if name == 'atom':
                dconv = mf.dm.cpu().numpy()
        d0 = guess(mol)
        S = None
        out['rms_learned'] = float(np.sqrt(np.mean((d0 - dconv)**2)))
        from gpu4pyscf_amd.scf.atom_guess import atom_guess
        out['rms_atom'] = float(np.sqrt(np.mean((atom_guess(mol) - dconv)**2)))
        rows.append(out)
        log(f"[eval {s}] cycles atom {out['atom']['cycles']} learned {out['learned']['cycles']}"
            f" dE {out['learned']['e'] - out['atom']['e']:.2e}")
    for k in ('atom', 'learned'):
        res[f'cycles_{k}'] = float(np.mean([r[k]['cycles'] for r in rows]))
        res[f'time_{k}_s'] = float(np.mean([r[k]['t'] for r in rows]))
        res[f'converged_{k}'] = all(r[k]['converged'] for r in rows)
    res['max_abs_dE'] = float(max(abs(r['learned']['e'] - r['atom']['e']) for r in rows))
    res['rms_guess_learned'] = float(np.mean([r['rms_learned'] for r in rows]))
    res['rms_guess_atom'] = float(np.mean([r['rms_atom'] for r in rows]))
    print(json.dumps(res), flush=True)


if __name__ == '__main__':
    main()
