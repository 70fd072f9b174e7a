'''Direct-SCF Coulomb (J) and exchange (K) matrices on MI355X.

Drop-in for gpu4pyscf.scf.jk (gpu4pyscf/scf/jk.py:60-230, 407-641): the same
``get_jk``/``get_j``/``get_k`` signatures and a ``_VHFOpt`` screening object
with the same attribute surface (``sorted_mol``, ``bas_pair_cache``,
``h_shls``, ``coeff``, ``apply_coeff_C_mat_CT``, ``apply_coeff_CT_mat_C``,
``get_jk(dms, hermi, verbose)``).  Device buffers are torch tensors; every
integral is evaluated by the HIP kernels of libjkhip.so.

Accepted ``dm``: numpy arrays (returned as numpy) or torch tensors (returned
as torch tensors on the same device).  Range-separated operators (omega,
lr_factor, sr_factor) follow gpu4pyscf/scf/jk.py:1122-1168.  Classes with
shells up to l = 3 run on the templated per-lane/team kernels, l = 4 and 5
on the runtime-l kernel; l > 5 raises NotImplementedError (the reference
sends l > 4 to PySCF on the CPU, gpu4pyscf/scf/jk.py:622-640; no CPU path
exists here).
'''

import ctypes
import math
import time

import numpy as np

from gpu4pyscf_amd.gto.sorted_gto import SortedGTO
from gpu4pyscf_amd.lib import clib

__all__ = ['get_jk', 'get_j', 'get_k', '_VHFOpt']

TILE = 12                 # pair enumeration tile, gpu4pyscf/scf/jk.py:51
KL_WINDOW_AO = 256        # widest kl column window (AOs of k, and of l), per-lane kernels
KL_WINDOW_AO_TEAM = 0     # the same for the team kernels (0: none; they run faster
                          # with K atomics than with LDS rows over windows)
Q_COND_MARGIN = 4.        # bin width of the descending q_cond ordering
LMAX = clib.JK_LMAX


def csr_rows(mat):
    '''(ptr, idx, coef) of the nonzeros of a dense host matrix, row by row
    (int32, int32, float64): row i holds idx[ptr[i]:ptr[i+1]] and their
    values.'''
    mat = np.asarray(mat, dtype=np.float64)
    r, c = np.nonzero(mat)
    ptr = np.searchsorted(r, np.arange(mat.shape[0] + 1)).astype(np.int32)
    return ptr, c.astype(np.int32), np.ascontiguousarray(mat[r, c])


def _ao_transform(mat, rows):
    '''out = A M A^T over the last two axes of a float64 device tensor, with A
    given by its compressed rows and its column count (jk_ao_transform).'''
    torch = _torch()
    ptr, idx, coef, n_in = rows
    n_out = ptr.numel() - 1
    if mat.dtype != torch.float64 or not mat.is_cuda:
        raise TypeError('AO transforms take float64 device tensors')
    if mat.dim() < 2 or tuple(mat.shape[-2:]) != (n_in, n_in):
        raise ValueError(f'AO transform of a {tuple(mat.shape)} tensor; '
                         f'the basis has {n_in} AOs')
    lead = tuple(mat.shape[:-2])
    src = mat.reshape((-1, n_in, n_in)).contiguous()
    out = torch.empty(src.shape[:1] + (n_out, n_out), dtype=torch.float64, device=mat.device)
    if src.shape[0]:
        clib.check(clib.load().jk_ao_transform(
            out.data_ptr(), src.data_ptr(), src.shape[0], n_out, n_in, ptr.data_ptr(),
            idx.data_ptr(), coef.data_ptr(), _stream_ptr()), 'jk_ao_transform')
    return out.reshape(lead + (n_out, n_out))


def _torch():
    import torch
    return torch


def _device():
    torch = _torch()
    if not torch.cuda.is_available():
        raise RuntimeError('the J/K engine needs a HIP device (MI355X); none is visible')
    return torch.device('cuda', torch.cuda.current_device())


def _stream_ptr():
    return ctypes.c_void_p(_torch().cuda.current_stream().cuda_stream)


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def tile_pairs(ish0, ish1, jsh0, jsh1, nbas, tile=TILE):
    '''Pair ids ish*nbas+jsh enumerated tile by tile (row-major inside a
    tile), the order of gpu4pyscf/lib/pbc/supmol_sr_estimator.cu:461-487.'''
    ni, nj = ish1 - ish0, jsh1 - jsh0
    if ni == 0 or nj == 0:
        return np.zeros(0, dtype=np.int64)
    ti = np.arange(ni) // tile
    tj = np.arange(nj) // tile
    ii, jj = np.meshgrid(np.arange(ni), np.arange(nj), indexing='ij')
    key = (ti[ii] * (nj // tile + 1) + tj[jj]) * (tile * tile) + (ii % tile) * tile + jj % tile
    order = np.argsort(key.ravel(), kind='stable')
    pairs = (ii.ravel() + ish0) * nbas + (jj.ravel() + jsh0)
    return pairs[order]


def group_by_split_points(q, split_points):
    '''Indices of ``q`` in descending 4-wide log bins; bin 0 (q <= log tol)
    is dropped (gpu4pyscf/pbc/scf/rsjk.py:2268-2277).'''
    bins = np.searchsorted(split_points, q.astype(np.float64))
    nb = len(split_points)
    subsets = [np.where(bins == i)[0] for i in range(nb, 0, -1)]
    if not subsets:
        return np.zeros(0, dtype=np.int64)
    return np.concatenate(subsets).astype(np.int64)


LANE_MAX = 170      # LANE_MAX of csrc/jk_kernels.hip
# class launches run on this many streams (32-water def2-TZVP build, ms:
# 1 -> 648, 2 -> 569, 3 -> 564, 4 -> 566 with four-wave per-lane blocks;
# 2 -> 482, 3 -> 463, 4 -> 454 with two-wave blocks).  HIP gives a process
# 4 hardware queues by default, so more streams would share them.
NSTREAMS = 4
# copies of vj/vk per build (32-water def2-TZVP build, ms, one stream:
# 1 -> 653, 2 -> 639, 4 -> 630, 8 -> 627, 16 -> 626; three streams:
# 1 -> 562, 8 -> 549)
NREPL = 8
REPL_BYTES = 16 << 30


# classes above LANE_MAX that the per-lane kernel evaluates in chunks of rows
# (LaneChunked of csrc/jk_class_kernels.h)
LANE_CHUNKED = {(3, 1, 2, 0), (3, 2, 1, 0), (2, 1, 1, 1), (2, 2, 2, 0), (3, 1, 1, 1),
                (3, 3, 1, 0), (3, 1, 3, 0)}


def class_kernel_name(ls):
    '''Name of the HIP kernel template that evaluates angular class ls =
    (li, lj, lk, ll): the per-lane kernel when the Cartesian block and the
    three 2D tables fit in LANE_MAX doubles or the class is one of
    LANE_CHUNKED, else the team kernel; classes with a shell of l >= 4 run
    on the runtime-l kernel.'''
    if max(ls) > 3:
        return f'jk_generic_kernel<{ls[0]}, {ls[1]}, {ls[2]}, {ls[3]}>'
    nc = [(l + 1) * (l + 2) // 2 for l in ls]
    nf = nc[0] * nc[1] * nc[2] * nc[3]
    gs = (ls[0] + 1) * (ls[1] + 1) * (ls[2] + 1) * (ls[3] + 1)
    lane = nf + 3 * gs <= LANE_MAX or tuple(ls) in LANE_CHUNKED
    kind = 'jk_lane_kernel' if lane else 'jk_team_kernel'
    return f'{kind}<{ls[0]}, {ls[1]}, {ls[2]}, {ls[3]}>'


def _set_kl_win(a, sets):
    '''kl column windows of a JKTask / JKBuildArgs (sets: an entry of
    _VHFOpt._kl_windows, [per-lane, team], or None).'''
    for n in range(2):
        w = sets[n] if sets is not None else None
        t = a.kl_win[n]
        if w is None:
            t.pairs = t.q = t.pdata = t.win = None
            t.n_win = t.width = 0
        else:
            t.pairs, t.q, t.pdata, t.win = (x.data_ptr() for x in w[:4])
            t.n_win, t.width = len(w[3]), w[4]


def kl_window_table(pairs, offs, ao_loc, nbas, k, l, W):
    '''kl column windows of the pair list ``pairs`` (ids ksh*nbas + lsh) of
    group pair (k, l), or None when neither group exceeds W AOs.  Each group
    is cut into windows of whole shells, as equal as they can be, of at most
    W AOs or one shell (the shells of a group have one size); returns (order, table, width): the stable
    reordering of the pairs window by window (each window keeps the list's
    descending q bins), the int32 (n_windows, 6) rows {first, end, k_ao0,
    k_nao, l_ao0, l_nao} of the non-empty windows (JKBuildArgs.kl_win) and
    the largest k_nao + l_nao.'''
    aol = np.asarray(ao_loc, dtype=np.int64)
    k_nao = int(aol[offs[k+1]] - aol[offs[k]])
    l_nao = int(aol[offs[l+1]] - aol[offs[l]])
    if len(pairs) == 0 or (k_nao <= W and l_nao <= W):
        return None

    def cut(g, nao):
        nsh = int(offs[g+1] - offs[g])
        spw = max(1, W // (nao // nsh))     # most shells per window
        spw = -(-nsh // -(-nsh // spw))     # the same number of windows, evened out
        return spw, -(-nsh // spw)
    spk, _ = cut(k, k_nao)
    spl, nwl = cut(l, l_nao)
    p = np.asarray(pairs).astype(np.int64)
    ksh, lsh = p // nbas, p % nbas
    key = ((ksh - offs[k]) // spk) * nwl + (lsh - offs[l]) // spl
    order = np.argsort(key, kind='stable')
    ids, first = np.unique(key[order], return_index=True)
    end = np.append(first[1:], len(p))
    a, b = np.divmod(ids, nwl)
    k0 = offs[k] + a * spk
    k1 = np.minimum(k0 + spk, offs[k+1])
    l0 = offs[l] + b * spl
    l1 = np.minimum(l0 + spl, offs[l+1])
    tab = np.stack([first, end, aol[k0], aol[k1] - aol[k0], aol[l0], aol[l1] - aol[l0]],
                   axis=1).astype(np.int32)
    return order, tab, int((tab[:, 3] + tab[:, 5]).max())


class _VHFOpt:
    '''Screening object of the direct J/K build (gpu4pyscf/scf/jk.py:407-486).'''

    def __init__(self, mol, direct_scf_tol=1e-13, tile=TILE):
        self.mol = mol
        self.direct_scf_tol = direct_scf_tol
        self.tile = tile
        self.h_shls = None
        self.sorted_mol = None
        self.bas_pair_cache = None
        self._dev = None
        self.timings = {}
        self.last_quartets = None
        # profiling mode: every get_jk records per-class HIP events and
        # surviving-quartet counters without synchronising; profile_results()
        # collects them
        self.profile = False
        self._profile_pending = []
        # class tasks only add into vj/vk with atomics, so they may run
        # concurrently: they are spread round-robin over nstreams streams
        # (default priority) that join the caller's stream at the end
        self.nstreams = NSTREAMS
        self._streams = {}
        # copies of vj/vk the class kernels add into (one per XCD)
        self.n_repl = NREPL
        # optional predicate on (li, lj, lk, ll): the class tasks to run
        self.task_filter = None
        # kl column windows: group pairs wider than kl_window_ao AOs in k or
        # l are split into windows of at most that many AOs each, so that a
        # work item (ij pair, window) keeps its K rows in LDS; the per-lane
        # kernels use kl_window_ao, the team kernels kl_window_ao_team (the
        # launcher decides per launch; kl_win_policy 1: always when they fit)
        # tasks run with bra and ket exchanged where that puts the larger
        # block on the bra side (JKBuildArgs.swap_ijkl)
        self.swap_tasks = True
        self.kl_window_ao = KL_WINDOW_AO
        self.kl_window_ao_team = KL_WINDOW_AO_TEAM
        self.kl_win_policy = 0

    def profile_results(self):
        '''Per class task: (summed kernel milliseconds, surviving quartets,
        launches) over the get_jk calls made since the last call, in
        profiling mode.'''
        torch = _torch()
        torch.cuda.synchronize()
        out = {}
        for tasks, events, counter in self._profile_pending:
            counts = counter.cpu().numpy().tolist()
            for t, (e0, e1), c in zip(tasks, events, counts):
                a = out.setdefault(t, [0., 0, 0])
                a[0] += e0.elapsed_time(e1)
                a[1] += int(c)
                a[2] += 1
        self._profile_pending = []
        return out

    def reset(self, mol):
        self.mol = mol
        self.sorted_mol = None
        self.bas_pair_cache = None
        self._dev = None

    # -- setup -------------------------------------------------------------
    def build(self, group_size=None, verbose=None):
        torch = _torch()
        t0 = time.perf_counter()
        lib = clib.load()
        dev = _device()
        mol = self.sorted_mol = self._make_sorted_mol()
        lmax = int(mol.uniq_l_ctr[:, 0].max())
        if lmax > LMAX:
            # the reference evaluates l > 4 shells with PySCF on the CPU
            # (gpu4pyscf/scf/jk.py:622-640); here every class runs on the GPU
            # up to l = 5 and higher l is refused
            raise NotImplementedError(f'l = {lmax} shells: the GPU path supports l <= {LMAX}')
        # screening q_cond of the molecule's operator (full, erf or erfc)
        q_omega, q_lr, q_sr = rsh_factors(self.mol)
        self.h_shls = []
        nbas = mol.nbas
        d = {}
        d['bas'] = torch.as_tensor(mol._bas.astype(np.int32), device=dev)
        d['env'] = torch.as_tensor(mol.env_kernel, device=dev)
        d['ao_loc'] = torch.as_tensor(mol.ao_loc.astype(np.int32), device=dev)
        d['coeff'] = torch.as_tensor(mol.coeff, device=dev)
        # compressed rows of C (density in) and of C^T (J/K out)
        for key, mat in (('c_rows', mol.coeff), ('ct_rows', mol.coeff.T)):
            d[key] = tuple(torch.as_tensor(x, device=dev) for x in csr_rows(mat)) + (mat.shape[1],)
        rys_buf, rys = clib.rys_table(dev)
        d['rys_buf'] = rys_buf
        self._rys = rys
        stream = _stream_ptr()

        # shell-pair overlap mask (float32 log bound > log tol^2), lower triangle
        exps, cs = mol.diffuse_pgto()
        exps_d = torch.as_tensor(exps.astype(np.float32), device=dev)
        logc_d = torch.as_tensor(np.log(np.abs(cs.astype(np.float32))), device=dev)
        coords_d = torch.as_tensor(mol.bas_coords(), device=dev)
        ls_d = torch.as_tensor(mol._bas[:, 1].astype(np.int32), device=dev)
        mask_d = torch.empty((nbas, nbas), dtype=torch.int8, device=dev)
        clib.check(lib.jk_ovlp_mask(_ptr(mask_d), _ptr(exps_d), _ptr(logc_d), _ptr(coords_d),
                                    _ptr(ls_d), ctypes.c_int(nbas),
                                    ctypes.c_float(math.log(self.direct_scf_tol**2)), stream),
                   'jk_ovlp_mask')
        mask = mask_d.cpu().numpy().astype(bool).ravel()

        # tile-ordered pair lists of each group pair, masked
        offs = mol.group_offsets
        ng = len(offs) - 1
        keys = [(i, j) for i in range(ng) for j in range(i + 1)]
        plist = []
        for i, j in keys:
            p = tile_pairs(offs[i], offs[i+1], offs[j], offs[j+1], nbas, self.tile)
            plist.append(p[mask[p]])
        allp = np.concatenate(plist).astype(np.uint32) if plist else np.zeros(0, np.uint32)
        q_all = np.zeros(len(allp), dtype=np.float32)
        if len(allp):
            allp_d = torch.as_tensor(allp.astype(np.int64), device=dev).to(torch.int32)
            q_d = torch.empty(len(allp), dtype=torch.float32, device=dev)
            clib.check(lib.jk_pair_qcond(_ptr(q_d), _ptr(allp_d), ctypes.c_int(len(allp)),
                                         _ptr(d['bas']), _ptr(d['env']), ctypes.c_int(nbas),
                                         rys, ctypes.c_double(abs(q_omega)),
                                         ctypes.c_double(q_lr), ctypes.c_double(q_sr), stream),
                       'jk_pair_qcond')
            q_all = q_d.cpu().numpy()
        split = np.arange(math.log(self.direct_scf_tol), 2., Q_COND_MARGIN)
        cache = {}
        dev_pairs, dev_q, dev_pd = {}, {}, {}
        pos = 0
        uniq = mol.uniq_l_ctr
        for (i, j), p in zip(keys, plist):
            q = q_all[pos:pos+len(p)]
            pos += len(p)
            idx = self._order_pairs(q, split)
            pij = p[idx].astype(np.uint32)
            qij = q[idx].astype(np.float32)
            cache[i, j] = (pij, qij, qij)
            pt = torch.as_tensor(pij.astype(np.int64), device=dev).to(torch.int32)
            dev_pairs[i, j] = pt
            dev_q[i, j] = torch.as_tensor(qij, device=dev)
            ip, jp = int(uniq[i, 1]), int(uniq[j, 1])
            pd = torch.empty((max(len(pij), 1), ip * jp, 6), dtype=torch.float64, device=dev)
            clib.check(lib.jk_pair_data(_ptr(pd), _ptr(pt), ctypes.c_int(len(pij)),
                                        ctypes.c_int(ip), ctypes.c_int(jp), _ptr(d['bas']),
                                        _ptr(d['env']), ctypes.c_int(nbas), stream),
                       'jk_pair_data')
            dev_pd[i, j] = pd
        self.bas_pair_cache = cache
        d['kl_win'] = self._kl_windows(cache, dev_pairs, dev_q, dev_pd)
        d['pairs'] = dev_pairs
        d['q'] = dev_q
        d['pdata'] = dev_pd
        d['task_array'] = self._task_array(d)
        self._dev = d
        self.device = dev
        self.timings['build'] = time.perf_counter() - t0
        return self

    def _make_sorted_mol(self):
        return SortedGTO(self.mol, diffuse_cutoff=0.3)

    def _kl_windows(self, cache, dev_pairs, dev_q, dev_pd):
        '''kl column windows of every group pair wider than the window in k
        or in l: {(k, l): [lane windows, team windows]}, each None or
        (pairs, q, pdata, table, width), the device copies of
        kl_window_table's reordering and window rows.'''
        torch = _torch()
        mol = self.sorted_mol
        out = {}
        for (k, l), (p, _, _) in cache.items():
            sets = []
            for W in (int(self.kl_window_ao), int(self.kl_window_ao_team)):
                w = (kl_window_table(p, mol.group_offsets, mol.ao_loc, mol.nbas, k, l, W)
                     if W > 0 else None)
                if w is None:
                    sets.append(None)
                    continue
                order, tab, width = w
                dev = dev_pairs[k, l].device
                o = torch.as_tensor(order, device=dev)
                sets.append((dev_pairs[k, l][o].contiguous(), dev_q[k, l][o].contiguous(),
                             dev_pd[k, l][o].contiguous(), torch.as_tensor(tab, device=dev),
                             width))
            if sets[0] is not None or sets[1] is not None:
                out[k, l] = sets
        return out

    def _order_pairs(self, q, split):
        '''Order of a group pair's pairs: descending 4-wide q_cond bins, bin 0
        dropped.'''
        return group_by_split_points(q, split)

    # -- AO transforms (gpu4pyscf/scf/jk.py:444-460) ---------------------
    @property
    def coeff(self):
        return self.sorted_mol.coeff

    def apply_coeff_C_mat_CT(self, mat):
        '''original AO -> sorted Cartesian AO: C M C^T over the last two axes
        (gpu4pyscf/gto/mole.py:906-921), by the sparse transform kernel.'''
        if self._dev is None:
            self.build()
        return _ao_transform(mat, self._dev['c_rows'])

    def apply_coeff_CT_mat_C(self, mat):
        '''sorted Cartesian AO -> original AO: C^T M C over the last two axes
        (gpu4pyscf/gto/mole.py:923-938), by the sparse transform kernel.'''
        if self._dev is None:
            self.build()
        return _ao_transform(mat, self._dev['ct_rows'])

    # -- J/K on the sorted basis (gpu4pyscf/scf/jk.py:492-641) -----------
    def tasks(self):
        '''Class tasks (i, j, k, l) over angular groups, reference order
        (gpu4pyscf/scf/jk.py:520-524); empty pair lists are skipped.'''
        mol = self.sorted_mol
        ng = len(mol.l_ctr_counts)
        ls = mol.uniq_l_ctr[:, 0]
        out = []
        for i in range(ng):
            for j in range(i + 1):
                for k in range(i + 1):
                    for l in range(k + 1):
                        if not (len(self.bas_pair_cache[i, j][0]) and
                                len(self.bas_pair_cache[k, l][0])):
                            continue
                        if self.task_filter is None or self.task_filter(
                                tuple(int(ls[x]) for x in (i, j, k, l))):
                            out.append((i, j, k, l))
        return out

    def _task_array(self, d):
        '''ctypes array of JKTask (include/jkhip.h) for self.tasks(), built
        once per geometry: the per-task launch arguments of jk_build_tasks.'''
        tasks = self.tasks()
        arr = (clib.JKTask * max(len(tasks), 1))()
        for t, task in enumerate(tasks):
            self._fill_task(arr[t], d, task)
        return arr

    def _fill_task(self, a, d, task):
        '''Per-task fields of a JKTask / JKBuildArgs for class task (i, j, k,
        l); with swap_tasks, the (xs|xp) tasks (li = lk, lj = 0, ll = 1; the
        two group pairs differ) run with bra and ket exchanged
        (swap_ijkl), which puts the p block on the bra side, where the
        per-lane kernels reduce J over the wavefront instead of adding it
        per lane (64-water build, ms on one stream: (ps|pp) 65.1 -> 26.1,
        (ds|dp) 41.5 -> 27.9, (fs|fp) 39.7 -> 35.4; the other classes with
        the larger block on the ket side measured slower swapped).'''
        mol = self.sorted_mol
        uniq = mol.uniq_l_ctr
        goff = mol.group_offsets
        ls = [int(uniq[x, 0]) for x in task]
        i, j, k, l = task
        swap = (self.swap_tasks and (i, j) != (k, l) and ls[0] == ls[2] and
                ls[1] == 0 and ls[3] == 1)
        if swap:
            i, j, k, l = k, l, i, j
        a.li, a.lj, a.lk, a.ll = (int(uniq[x, 0]) for x in (i, j, k, l))
        a.iprim, a.jprim, a.kprim, a.lprim = (int(uniq[x, 1]) for x in (i, j, k, l))
        a.k_ao0 = int(mol.ao_loc[goff[k]])
        a.k_nao = int(mol.ao_loc[goff[k + 1]]) - a.k_ao0
        a.l_ao0 = int(mol.ao_loc[goff[l]])
        a.l_nao = int(mol.ao_loc[goff[l + 1]]) - a.l_ao0
        a.npairs_ij = len(d['pairs'][i, j])
        a.npairs_kl = len(d['pairs'][k, l])
        a.pairs_ij = d['pairs'][i, j].data_ptr()
        a.pairs_kl = d['pairs'][k, l].data_ptr()
        a.q_ij = d['q'][i, j].data_ptr()
        a.q_kl = d['q'][k, l].data_ptr()
        a.pdata_ij = d['pdata'][i, j].data_ptr()
        a.pdata_kl = d['pdata'][k, l].data_ptr()
        _set_kl_win(a, d['kl_win'].get((k, l)))
        a.swap_ijkl = int(swap)

    def operator_passes(self, omega=None, lr_factor=None, sr_factor=None):
        """Launch passes [(omega >= 0, scale)] of the operator
        lr_factor erf(|omega| r)/r + sr_factor erfc(|omega| r)/r
        = sr_factor / r + (lr_factor - sr_factor) erf(|omega| r)/r;
        omega=None takes the molecule's omega with the libcint sign
        convention (gpu4pyscf/scf/jk.py:1122-1168)."""
        omega, lr, sr = rsh_factors(self.mol, omega, lr_factor, sr_factor)
        if omega == 0:
            return [(0., float(lr))]
        passes = []
        if sr != 0:
            passes.append((0., float(sr)))
        if lr - sr != 0:
            passes.append((abs(float(omega)), float(lr - sr)))
        return passes

    def get_jk(self, dms, hermi, verbose=None, with_j=True, with_k=True,
               count_quartets=False, time_classes=False, fp32_tol=None,
               hash_quartets=False, omega=None, lr_factor=None, sr_factor=None,
               screen=None):
        """J/K for density matrices ``dms`` (n_dm, nao_c, nao_c) in the sorted
        Cartesian basis (torch, float64, on the device).

        count_quartets: store the surviving quartets of each class task in
        ``self.last_task_counts`` (and their sum in ``self.last_quartets``).
        hash_quartets: also store the set checksum of each task's surviving
        quartets (sum mod 2**64 of the SplitMix64 finaliser of
        bas_ij << 32 | bas_kl) in ``self.last_task_hashes``.
        time_classes: store (task, milliseconds) of each class launch, timed
        with HIP events on the stream it runs on, in ``self.last_task_ms``.
        fp32_tol: evaluate in FP32 the quartets whose bound
        sqrt|(ij|ij)| sqrt|(kl|kl)| max|D| is below fp32_tol (None: all FP64).
        omega, lr_factor, sr_factor: the operator of K (default: the
        molecule's omega); a range-separated operator takes two launches of
        every class task (full range and erf-attenuated), whose counters are
        those of the first.  J always uses the molecule's operator; when K's
        differs, J comes from a J-only build and K from a K-only build, as in
        gpu4pyscf/scf/jk.py:129-135 (the counters are then the K build's).
        screen: the density blocks of the survival test, 'jk' (all six, the
        default of a J/K build, RYS_build_jk), 'j' (D_ij, D_kl: a J-only
        build, RYS_build_j, whose densities are first symmetrised as in
        gpu4pyscf/scf/jk.py:224-226) or 'k' (D_ik, D_jk, D_il, D_jl: a
        K-only build, RYS_build_k)."""
        torch = _torch()
        if self._dev is None:
            self.build()
        if callable(dms):
            dms = dms()
        if (screen is None and omega is not None and with_j and with_k and
                self.operator_passes() != self.operator_passes(omega, lr_factor, sr_factor)):
            kw = dict(count_quartets=count_quartets, time_classes=time_classes,
                      fp32_tol=fp32_tol, hash_quartets=hash_quartets)
            vj = self.get_jk(dms, hermi, verbose, True, False, screen='j', **kw)[0]
            vk = self.get_jk(dms, hermi, verbose, False, True, omega=omega,
                             lr_factor=lr_factor, sr_factor=sr_factor, screen='k', **kw)[1]
            return vj, vk
        screen = screen or 'jk'
        if screen not in clib.JK_SCREEN or (screen == 'j' and with_k) or \
                (screen == 'k' and with_j):
            raise ValueError(f'screen {screen!r} with with_j={with_j}, with_k={with_k}')
        if screen == 'j' and hermi != 1:
            dms = (dms + dms.transpose(1, 2)) * .5
            hermi = 1
        lib = clib.load()
        mol = self.sorted_mol
        d = self._dev
        nbas, nao = mol.nbas, mol.nao
        assert dms.ndim == 3 and dms.shape[-1] == nao
        dms = dms.contiguous()
        stream = _stream_ptr()
        n_dm0 = dms.shape[0]
        dm_cond = torch.empty((nbas, nbas), dtype=torch.float32, device=dms.device)
        work = torch.empty((nbas, nbas), dtype=torch.float64, device=dms.device)
        clib.check(lib.jk_dm_cond(_ptr(dm_cond), _ptr(work), _ptr(dms), ctypes.c_int(n_dm0),
                                  ctypes.c_int(nao), _ptr(d['ao_loc']), ctypes.c_int(nbas),
                                  ctypes.c_int(hermi), stream), 'jk_dm_cond')
        dm_penalty = float(dm_cond.max())
        if hermi == 0:
            dms = torch.cat([dms, dms.transpose(1, 2)]).contiguous()
        n_dm = dms.shape[0]
        # n_repl zeroed copies of vj/vk; workgroup b adds into copy
        # b % n_repl (one per XCD), the copies are summed after the build
        n_repl = max(int(self.n_repl), 1)
        # at most REPL_BYTES for the copies of J and K together
        n_repl = max(1, min(n_repl, REPL_BYTES // max(2 * dms.numel() * 8, 1)))
        shape = (n_repl,) + tuple(dms.shape)
        vj = torch.zeros(shape, dtype=dms.dtype, device=dms.device) if with_j else None
        vk = torch.zeros(shape, dtype=dms.dtype, device=dms.device) if with_k else None
        tasks = self.tasks()
        prof = self.profile
        count_quartets = count_quartets or prof
        counter = (torch.zeros(len(tasks), dtype=torch.int64, device=dms.device)
                   if count_quartets else None)
        counter32 = (torch.zeros(len(tasks), dtype=torch.int64, device=dms.device)
                     if count_quartets else None)
        hashes = (torch.zeros(len(tasks), dtype=torch.int64, device=dms.device)
                  if hash_quartets else None)
        events = []

        args = clib.JKBuildArgs()
        args.abi_version = clib.JK_ABI_VERSION
        args.nbas = nbas
        args.nao = nao
        args.bas = d['bas'].data_ptr()
        args.env = d['env'].data_ptr()
        args.ao_loc = d['ao_loc'].data_ptr()
        args.dm_cond = dm_cond.data_ptr()
        args.cutoff = np.float32(math.log(self.direct_scf_tol))
        args.dm_penalty = dm_penalty
        args.n_dm = n_dm
        args.dm = dms.data_ptr()
        args.rys = self._rys
        args.fp32_cut = np.float32(math.log(fp32_tol)) if fp32_tol else np.float32(-1e30)
        args.n_repl = n_repl
        args.repl_stride = dms.numel()
        args.screen = clib.JK_SCREEN[screen]
        args.kl_win_policy = int(self.kl_win_policy)
        # launch passes of the operator: (omega, scale, with J, with K)
        runs = [(w, f, with_j, with_k)
                for w, f in self.operator_passes(omega, lr_factor, sr_factor)]
        cur = torch.cuda.current_stream(dms.device)
        streams = [cur]
        if self.nstreams > 1:
            key = str(dms.device)
            if key not in self._streams:
                self._streams[key] = [torch.cuda.Stream(dms.device)
                                      for _ in range(self.nstreams - 1)]
            streams += self._streams[key][:self.nstreams - 1]
            ready = torch.cuda.Event()
            ready.record(cur)
            for s in streams[1:]:
                s.wait_event(ready)
        # the side streams are joined back into the caller's stream even when a
        # launch fails, so that no kernel still adds into vj/vk after they are
        # released
        try:
            for n_run, (w, f, rj, rk) in enumerate(runs):
                args.omega = w
                args.scale = f
                args.vj = vj.data_ptr() if rj else None
                args.vk = vk.data_ptr() if rk else None
                first = n_run == 0
                # launched from the last task (f classes, the longest) to the
                # first, so that the short s/p classes fill the streams at the
                # end (32-water def2-TZVP build: 521 ms in task order, 515
                # reversed).  Without per-launch timing the whole loop is one
                # native call.
                if not (time_classes or prof):
                    args.n_quartets = counter.data_ptr() if (first and counter is not None) else None
                    args.n_quartets_fp32 = (counter32.data_ptr()
                                            if (first and counter is not None) else None)
                    args.quartet_hash = hashes.data_ptr() if (first and hashes is not None) else None
                    handles = (ctypes.c_void_p * len(streams))(*[s.cuda_stream for s in streams])
                    clib.check(lib.jk_build_tasks(ctypes.byref(args), d['task_array'], len(tasks),
                                                  handles, len(streams)), 'jk_build_tasks')
                    continue
                for n, t in enumerate(range(len(tasks) - 1, -1, -1)):
                    task_stream = streams[n % len(streams)]
                    args.n_quartets = (counter.data_ptr() + 8 * t
                                       if (first and counter is not None) else None)
                    args.n_quartets_fp32 = (counter32.data_ptr() + 8 * t
                                            if (first and counter is not None) else None)
                    args.quartet_hash = (hashes.data_ptr() + 8 * t
                                         if (first and hashes is not None) else None)
                    self._fill_task(args, d, tasks[t])
                    if first and (time_classes or prof):
                        e0 = torch.cuda.Event(enable_timing=True)
                        e1 = torch.cuda.Event(enable_timing=True)
                        e0.record(task_stream)
                    st = lib.jk_build(ctypes.byref(args), ctypes.c_void_p(task_stream.cuda_stream))
                    if first and (time_classes or prof):
                        e1.record(task_stream)
                        events.append((t, e0, e1))
                    if st != 0:
                        clib.check(st, f'jk_build ({"spdfghi"[args.li]}{"spdfghi"[args.lj]}|'
                                       f'{"spdfghi"[args.lk]}{"spdfghi"[args.ll]})')
        finally:
            for s in streams[1:]:
                done = torch.cuda.Event()
                done.record(s)
                cur.wait_event(done)
        if with_j:
            vj = vj[0] if n_repl == 1 else vj.sum(0)
        if with_k:
            vk = vk[0] if n_repl == 1 else vk.sum(0)
        if prof:
            events.sort(key=lambda e: e[0])
            self._profile_pending.append((tasks, [(e0, e1) for _, e0, e1 in events], counter))
        elif counter is not None:
            self.last_task_counts = dict(zip(tasks, counter.cpu().numpy().tolist()))
            self.last_quartets = int(sum(self.last_task_counts.values()))
            self.last_quartets_fp32 = int(counter32.sum().item())
        if hashes is not None:
            self.last_task_hashes = dict(zip(tasks, (int(h) & (2**64 - 1)
                                                     for h in hashes.cpu().numpy().tolist())))
        if time_classes and not prof:
            torch.cuda.synchronize()
            self.last_task_ms = [(tasks[t], e0.elapsed_time(e1)) for t, e0, e1 in sorted(events)]
        if hermi == 1:
            if with_j:
                vj *= 2.
            if with_k:
                vk = vk + vk.transpose(1, 2)
        elif hermi == 2:
            if with_j:
                vj.zero_()
            if with_k:
                vk = vk - vk.transpose(1, 2)
        else:
            h = n_dm // 2
            if with_j:
                vj = vj[:h] + vj[h:].transpose(1, 2)
            if with_k:
                vk = vk[:h] + vk[h:].transpose(1, 2)
        if with_j:
            vj = vj + vj.transpose(1, 2)
        return vj, vk


def rsh_factors(mol, omega=None, lr_factor=None, sr_factor=None):
    """(omega, lr_factor, sr_factor) of the operator
    lr_factor erf(|omega| r)/r + sr_factor erfc(|omega| r)/r, following
    _check_rsh_factors (gpu4pyscf/scf/jk.py:1122-1168): omega None takes
    mol.omega; without factors omega > 0 is long range, omega < 0 short
    range."""
    if omega is None:
        omega = getattr(mol, 'omega', 0.) or 0.
    if lr_factor is None and sr_factor is None:
        if omega == 0:
            lr_factor = sr_factor = 1
        elif omega < 0:
            lr_factor, sr_factor = 0, 1
        else:
            lr_factor, sr_factor = 1, 0
    elif lr_factor is None:
        lr_factor = sr_factor if omega == 0 else 0
    elif sr_factor is None:
        if omega == 0:
            sr_factor = lr_factor
        else:
            assert omega > 0
            sr_factor = 0
    else:
        if omega == 0:
            assert lr_factor == sr_factor
        elif lr_factor == sr_factor:
            omega = 0
    if sr_factor != 0 and omega != 0:
        omega = -abs(omega)
    return omega, lr_factor, sr_factor


def _to_device(dm):
    torch = _torch()
    if isinstance(dm, torch.Tensor):
        return dm.to(device=_device(), dtype=torch.float64), True
    return torch.as_tensor(np.asarray(dm, dtype=np.float64), device=_device()), False


def get_jk(mol, dm, hermi=0, vhfopt=None, with_j=True, with_k=True,
           omega=None, lr_factor=None, sr_factor=None, verbose=None):
    """Coulomb and exchange matrices, gpu4pyscf/scf/jk.py:60-151.

    Returns (vj, vk) with the shape of ``dm``; an element is None when not
    requested.  With omega (and lr_factor, sr_factor) given, K uses the
    operator lr_factor erf(|omega| r)/r + sr_factor erfc(|omega| r)/r and J
    the molecule's operator (mol.omega), as in the reference.  The quartets
    are screened on all six density blocks (the reference's RYS_build_jk)
    whichever matrices are requested."""
    return _get_jk(mol, dm, hermi, vhfopt, with_j, with_k, omega, lr_factor, sr_factor, None)


def _get_jk(mol, dm, hermi, vhfopt, with_j, with_k, omega, lr_factor, sr_factor, screen):
    assert with_j or with_k
    if vhfopt is None:
        vhfopt = _VHFOpt(mol).build()
    dm_t, was_torch = _to_device(dm)
    shape = dm_t.shape
    nao0 = shape[-1]
    dms = dm_t.reshape(-1, nao0, nao0)
    dms = vhfopt.apply_coeff_C_mat_CT(dms)
    vj, vk = vhfopt.get_jk(dms, hermi, with_j=with_j, with_k=with_k,
                           omega=omega, lr_factor=lr_factor, sr_factor=sr_factor, screen=screen)
    if with_j:
        vj = vhfopt.apply_coeff_CT_mat_C(vj).reshape(shape)
    if with_k:
        vk = vhfopt.apply_coeff_CT_mat_C(vk).reshape(shape)
    if not was_torch:
        vj = vj.cpu().numpy() if with_j else None
        vk = vk.cpu().numpy() if with_k else None
    return vj, vk


def get_j(mol, dm, hermi=0, vhfopt=None, verbose=None):
    '''Coulomb matrix by a J-only build (screened on D_ij and D_kl),
    gpu4pyscf/scf/jk.py:206-230 and its _VHFOpt.get_j (:643-783).'''
    return _get_jk(mol, dm, hermi, vhfopt, True, False, None, None, None, 'j')[0]


def get_k(mol, dm, hermi=0, vhfopt=None, omega=None, lr_factor=None, sr_factor=None):
    '''Exchange matrix by a K-only build (screened on D_ik, D_jk, D_il and
    D_jl), gpu4pyscf/scf/jk.py:153-204 and its _VHFOpt.get_k (:785-927).'''
    return _get_jk(mol, dm, hermi, vhfopt, False, True, omega, lr_factor, sr_factor, 'k')[1]
