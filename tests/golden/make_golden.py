"""Write golden vectors for sle.als / evp.als with the reference itself.

Run in the build container only (the reference is not available elsewhere):

    PYTHONDONTWRITEBYTECODE=1 python tests/golden/make_golden.py /root/reference

Only the two solvers are recorded here: their exact iterates (which local
solution, which orthonormal factor, best-so-far tracking, the convergence
rule and the iteration count) are pinned by nothing but the reference.  The
TT algebra, stacks, micro matrices and local solvers are pinned by dense
brute force and closed forms in tests/test_oracle.py instead.

Each case stores its inputs (so the fixture does not depend on this repo's
input builders), the returned cores and ranks, and gauge-invariant scalars.
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))


def _save_tt(store, prefix, cores):
    store[prefix + '_order'] = np.array(len(cores))
    for i, c in enumerate(cores):
        store['%s_%d' % (prefix, i)] = np.ascontiguousarray(c)


def main(ref_root):
    sys.path.insert(0, ref_root)
    sys.path.insert(0, ROOT)
    from scikit_tt.tensor_train import TT
    import scikit_tt.tensor_train as rtt
    import scikit_tt.solvers.sle as rsle
    import scikit_tt.solvers.evp as revp
    from scikit_tt_amd import workloads as W

    out = {}

    # 1. sle.als, -Laplacian d=10 n=8, b = ones, x0 N(0,1) rank 5 (rng 0), 4 sweeps, 'solve'
    A, x0, b = W.laplace_small()
    x = rsle.als(TT(A), TT(x0), TT(b), repeats=4, solver='solve')
    s = {}
    _save_tt(s, 'A', A)
    _save_tt(s, 'x0', x0)
    _save_tt(s, 'b', b)
    _save_tt(s, 'x', x.cores)
    s['ranks'] = np.array(x.ranks)
    s['relres'] = np.array((TT(A).dot(x) - TT(b)).norm() / TT(b).norm())
    s['norm'] = np.array(x.norm())
    np.savez_compressed(os.path.join(HERE, 'sle_laplace_d10.npz'), **s)
    out['sle_laplace_d10'] = float(s['relres'])

    # 2. sle.als, indefinite Toeplitz operator (reference test pattern,
    #    tests/test_sle.py:20-35, made smaller): toeplitz(1..64) as a TT
    #    operator with [2]*6 modes, truncated TT-SVD; x0 = ones rank 3,
    #    ortho_right; 2 sweeps, 'lu'
    import scipy.linalg as sla
    mat = sla.toeplitz(np.arange(1, 65), np.arange(1, 65)).astype(float)
    op = TT(mat.reshape([2] * 12), threshold=1e-12)
    x0t = rtt.ones(op.row_dims, [1] * 6, ranks=3).ortho_right()
    bt = rtt.ones(op.row_dims, [1] * 6)
    xt = rsle.als(op, x0t, bt, repeats=2, solver='lu')
    s = {}
    _save_tt(s, 'A', op.cores)
    _save_tt(s, 'x0', x0t.cores)
    _save_tt(s, 'b', bt.cores)
    _save_tt(s, 'x', xt.cores)
    s['ranks'] = np.array(xt.ranks)
    s['dense_solution'] = np.linalg.solve(mat, np.ones(64))
    np.savez_compressed(os.path.join(HERE, 'sle_toeplitz_d6.npz'), **s)

    # 3. evp.als 'eig' sigma=-100 (lowest), TFIM d=8 rank 8, repeats=8,
    #    conv_eps=1e-12 (the convergence rule may stop early)
    for tag, d, r, reps, eps, solver, sigma in [('evp_tfim_d8_eig', 8, 8, 8, 1e-12, 'eig', -100.0),
                                                ('evp_tfim_d8_eigh', 8, 6, 5, 0.0, 'eigh', 1.0),
                                                ('evp_tfim_d12_conv', 12, 6, 10, 1e-9, 'eig', -100.0)]:
        H, y0, J, g = W.tfim_chain(seed=3, d=d, rank=r)
        y0t = TT(y0).ortho_right()
        y0t = (1.0 / y0t.norm()) * y0t
        lam, y, it = revp.als(TT(H), y0t, repeats=reps, conv_eps=eps, solver=solver, sigma=sigma)
        s = {}
        _save_tt(s, 'A', H)
        _save_tt(s, 'x0', y0t.cores)
        _save_tt(s, 'x', y.cores)
        s['ranks'] = np.array(y.ranks)
        s['eigenvalue'] = np.array(float(np.real(lam)))
        s['iterations'] = np.array(it)
        s['repeats'] = np.array(reps)
        s['conv_eps'] = np.array(eps)
        s['sigma'] = np.array(sigma)
        s['solver'] = np.array(solver)
        s['J'] = J
        s['g'] = g
        np.savez_compressed(os.path.join(HERE, tag + '.npz'), **s)
        out[tag] = (float(np.real(lam)), int(it))
    print(out)


def c2_instance(ref_root):
    """evp.als on the benchmark's TFIM instance 0 (d=32, n=2, MPO rank 3,
    rank-16 N(0,1) start right-orthonormalised and normalised), 6 sweeps,
    'eig', sigma=-100, conv_eps=0: the full-size eigenproblem configuration."""
    sys.path.insert(0, ref_root)
    sys.path.insert(0, ROOT)
    from scikit_tt.tensor_train import TT
    import scikit_tt.solvers.evp as revp
    from scikit_tt_amd import workloads as W

    H, y0, J, g = W.tfim_chain(seed=0, d=32, rank=16)
    y0t = TT(y0).ortho_right()
    y0t = (1.0 / y0t.norm()) * y0t
    lam, y, it = revp.als(TT(H), y0t, repeats=6, conv_eps=0, solver='eig', sigma=-100.0)
    s = {}
    _save_tt(s, 'A', H)
    _save_tt(s, 'x0', y0t.cores)
    _save_tt(s, 'x', y.cores)
    s['ranks'] = np.array(y.ranks)
    s['eigenvalue'] = np.array(float(np.real(lam)))
    s['iterations'] = np.array(it)
    s['repeats'] = np.array(6)
    s['conv_eps'] = np.array(0.0)
    s['sigma'] = np.array(-100.0)
    s['solver'] = np.array('eig')
    s['J'] = J
    s['g'] = g
    np.savez_compressed(os.path.join(HERE, 'evp_tfim_d32_c2.npz'), **s)
    print('evp_tfim_d32_c2', float(np.real(lam)), int(it))


def ode_cases(ref_root):
    """Implicit integrators on the reference's own ODE test operator
    (signaling cascade, 2 species, quantised to 12 modes of size 2,
    tests/test_ode.py:176-192) with fewer steps: implicit Euler (20 unit
    steps), trapezoidal rule (mixed steps) and the adaptive method; the
    operator, start values and the returned trains are stored."""
    sys.path.insert(0, ref_root)
    sys.path.insert(0, ROOT)
    import scikit_tt.tensor_train as rtt
    import scikit_tt.models as mdl
    import scikit_tt.solvers.ode as rode
    op = mdl.signaling_cascade(2).tt2qtt([[2] * 6] * 2, [[2] * 6] * 2)
    iv = rtt.unit(op.row_dims, [0] * op.order)
    ig = rtt.ones(op.row_dims, [1] * op.order, ranks=4).ortho_right()
    s = {}
    _save_tt(s, 'A', op.cores)
    _save_tt(s, 'iv', iv.cores)
    _save_tt(s, 'ig', ig.cores)
    ie = rode.implicit_euler(op, iv, ig, [1.0] * 20, progress=False)
    _save_tt(s, 'ie_last', ie[-1].cores)
    s['ie_errors'] = np.array(rode.errors_impl_euler(op, ie, [1.0] * 20))
    steps = [0.01] * 5 + [0.5] * 5
    tr = rode.trapezoidal_rule(op, iv, ig, steps, progress=False)
    _save_tt(s, 'tr_last', tr[-1].cores)
    s['tr_errors'] = np.array(rode.errors_trapezoidal(op, tr, steps))
    ad, times = rode.adaptive_step_size(op, iv, ig, 30, step_size_first=1, second_method='two_step_Euler',
                                        progress=False)
    _save_tt(s, 'ad_last', ad[-1].cores)
    s['ad_times'] = np.array(times)
    np.savez_compressed(os.path.join(HERE, 'ode_signaling_cascade.npz'), **s)
    print('ode', s['ie_errors'][-1], s['tr_errors'][-1], len(times))


def power_case(ref_root):
    """evp.power_method on a TFIM chain (d=6, rank 4 start), sigma below the
    spectrum: the inverse iteration converges to the ground state."""
    sys.path.insert(0, ref_root)
    sys.path.insert(0, ROOT)
    from scikit_tt.tensor_train import TT
    import scikit_tt.solvers.evp as revp
    from scikit_tt_amd import workloads as W
    H, y0, J, g = W.tfim_chain(seed=5, d=6, rank=4)
    y0t = TT(y0).ortho_right()
    y0t = (1.0 / y0t.norm()) * y0t
    lam, y = revp.power_method(TT(H), y0t, repeats=8, sigma=-9.0)
    s = {}
    _save_tt(s, 'A', H)
    _save_tt(s, 'x0', y0t.cores)
    _save_tt(s, 'x', y.cores)
    s['eigenvalue'] = np.array(float(lam))
    s['sigma'] = np.array(-9.0)
    s['repeats'] = np.array(8)
    s['exact'] = np.array(W.tfim_ground_energy(J, g))
    np.savez_compressed(os.path.join(HERE, 'evp_power_tfim_d6.npz'), **s)
    print('power', float(lam), W.tfim_ground_energy(J, g))


def mals_case(ref_root):
    """sle.mals on the reference's own test system (tests/test_sle.py:14-35,
    61-85): the Toeplitz matrix toeplitz(1..1024) as a TT operator (TT-SVD of
    the full array at threshold 1e-12), all-ones rhs, ones rank-5 start right-orthonormalised,
    one sweep, threshold 1e-14, max_rank 10, for 'solve' and 'lu'.  Also the
    TT-SVD of that array (operator cores and ranks) and tt2qtt of a small
    train."""
    sys.path.insert(0, ref_root)
    sys.path.insert(0, ROOT)
    import scipy.linalg as sla
    from scikit_tt.tensor_train import TT
    import scikit_tt.tensor_train as rtt
    import scikit_tt.solvers.sle as rsle
    d = 10
    mat = sla.toeplitz(np.arange(1, 2 ** d + 1), np.arange(1, 2 ** d + 1)).astype(float)
    # (truncated at 1e-12: the untruncated TT-SVD of the test keeps numerically zero singular
    # values and ranks up to 2^d; the local problems of (M)ALS depend only on the operator itself)
    op = TT(mat.reshape([2] * 2 * d), threshold=1e-12)
    rhs = rtt.ones(op.row_dims, [1] * d)
    x0 = rtt.ones(op.row_dims, [1] * d, ranks=5).ortho_right()
    s = {}
    _save_tt(s, 'A', op.cores)
    s['A_ranks'] = np.array(op.ranks)
    _save_tt(s, 'x0', x0.cores)
    _save_tt(s, 'b', rhs.cores)
    for solver in ('solve', 'lu'):
        x = rsle.mals(op, x0, rhs, repeats=1, solver=solver, threshold=1e-14, max_rank=10)
        _save_tt(s, 'x_' + solver, x.cores)
        s['ranks_' + solver] = np.array(x.ranks)
    s['dense_solution'] = np.linalg.solve(mat, np.ones(2 ** d))
    np.savez_compressed(os.path.join(HERE, 'sle_mals_toeplitz_d10.npz'), **s)
    # TT-SVD with truncation of a smaller full operator and tt2qtt of a train
    rng = np.random.default_rng(9)
    small = sla.toeplitz(np.arange(1, 65), np.arange(1, 65)).astype(float) + 0.01 * rng.standard_normal((64, 64))
    t = {}
    for thr, tag in ((0.0, 'exact'), (1e-10, 'thr')):
        tt_ = TT(small.reshape([2] * 12), threshold=thr)
        _save_tt(t, 'svd_' + tag, tt_.cores)
        t['svd_ranks_' + tag] = np.array(tt_.ranks)
    tt_ = TT(small.reshape([4] * 6), max_rank=5)
    _save_tt(t, 'svd_maxr', tt_.cores)
    t['svd_ranks_maxr'] = np.array(tt_.ranks)
    t['full'] = small
    base = TT([rng.standard_normal((1, 4, 4, 3)), rng.standard_normal((3, 8, 2, 2)), rng.standard_normal((2, 4, 4, 1))])
    _save_tt(t, 'q_in', base.cores)
    q = base.tt2qtt([[2, 2], [2, 2, 2], [4]], [[2, 2], [2, 1, 1], [2, 2]])
    _save_tt(t, 'q_out', q.cores)
    t['q_ranks'] = np.array(q.ranks)
    qt = base.tt2qtt([[2, 2], [2, 2, 2], [4]], [[2, 2], [2, 1, 1], [2, 2]], threshold=1e-2)
    _save_tt(t, 'q_out_thr', qt.cores)
    t['q_ranks_thr'] = np.array(qt.ranks)
    np.savez_compressed(os.path.join(HERE, 'tt_svd_qtt.npz'), **t)
    print('mals ranks', list(s['ranks_solve']), list(s['ranks_lu']), 'svd ranks', list(t['svd_ranks_exact']),
          list(t['svd_ranks_thr']), list(t['svd_ranks_maxr']), 'qtt', list(t['q_ranks']), list(t['q_ranks_thr']))


def evp_options_case(ref_root):
    """evp.als with the options that run outside the fused engine.

    * The reference's own evp test problem (tests/test_evp.py): the Ulam
      operator of the triple-well transitions (examples/data/
      triple_well_transitions.npz, coarse-grained by 2, 25 x 25 states, 2000
      simulations), ones rank-15 start right-orthonormalised, 10 sweeps:
      'eig' with operator_gevp = identity and with number_ev = 3 on the
      operator as it is (not symmetric; test_als_eig), and number_ev = 3
      'eigh' on its symmetric part (test_als_eigh).
    * A TFIM chain (d = 6, rank-4 start): a generalised problem with a
      symmetric positive definite operator_gevp (a rank-1 TT of positive
      diagonals), 'eig' nearest sigma = -100; deflation of the ground state
      (previous, shift = 50) from another start; number_ev = 2 with 'eig'.
    Dense eigenvalues of every problem are stored beside the reference's."""
    sys.path.insert(0, ref_root)
    sys.path.insert(0, ROOT)
    import scipy.linalg as sla
    from scikit_tt.tensor_train import TT
    import scikit_tt.tensor_train as rtt
    import scikit_tt.solvers.evp as revp
    import scikit_tt.data_driven.ulam as ulam
    from scikit_tt_amd import workloads as W
    s = {}
    transitions = np.load(os.path.join(ref_root, 'examples', 'data', 'triple_well_transitions.npz'))['transitions']
    transitions = np.int64(np.ceil(np.true_divide(transitions, 2)))
    op = ulam.ulam_2d(transitions, [25, 25], 2000)
    x0 = rtt.ones(op.row_dims, [1] * op.order, ranks=15).ortho_right()
    # the reference test's 'eig' cases on the operator as it is (not symmetric):
    # one eigenpair with operator_gevp = identity, and three
    lam, vec, it = revp.als(op, x0, operator_gevp=rtt.eye(op.row_dims), repeats=10)
    _save_tt(s, 'ulam_ns_A', op.cores)
    s['ulam_ns_val'] = np.array(float(np.real(lam)))
    s['ulam_ns_iters'] = np.array(it)
    _save_tt(s, 'ulam_ns_vec', [np.real(c) for c in vec.cores])
    vals, vecs, it = revp.als(op, x0, number_ev=3, repeats=10)
    s['ulam_ns3_vals'] = np.real(np.array(vals, dtype=complex))
    s['ulam_ns3_iters'] = np.array(it)
    for e in range(3):
        _save_tt(s, 'ulam_ns3_vec%d' % e, [np.real(c) for c in vecs[e].cores])
    dense_ns = np.linalg.eigvals(op.matricize())
    s['ulam_ns_dense_top3'] = np.real(dense_ns[np.argsort(-np.real(dense_ns))][:3])
    s['ulam_ns_complex_parts'] = np.array(max(float(np.max(np.abs(np.imag(c)))) for c in vecs[0].cores))
    op = 0.5 * (op + op.transpose())
    vals, vecs, it = revp.als(op, x0, number_ev=3, repeats=10, solver='eigh')
    _save_tt(s, 'ulam_A', op.cores)
    _save_tt(s, 'ulam_x0', x0.cores)
    s['ulam_vals'] = np.array(vals, dtype=float)
    s['ulam_iters'] = np.array(it)
    for e in range(3):
        _save_tt(s, 'ulam_vec%d' % e, vecs[e].cores)
    s['ulam_ranks'] = np.array(vecs[0].ranks)
    dense = np.linalg.eigvalsh(op.matricize())
    s['ulam_dense_top3'] = dense[::-1][:3].copy()
    # TFIM chain
    H, y0, J, g = W.tfim_chain(seed=21, d=6, rank=4)
    rng = np.random.default_rng(22)
    y0t = TT(y0).ortho_right()
    y0t = (1.0 / y0t.norm()) * y0t
    y1 = TT(W.random_tt(rng, [2] * 6, 4)).ortho_right()
    y1 = (1.0 / y1.norm()) * y1
    Mcores = [np.diag(1.0 + 0.3 * rng.random(2)).reshape(1, 2, 2, 1) for _ in range(6)]
    _save_tt(s, 'tfim_A', H)
    _save_tt(s, 'tfim_x0', y0t.cores)
    _save_tt(s, 'tfim_x1', y1.cores)
    _save_tt(s, 'tfim_M', Mcores)
    Hm = TT(H).matricize()
    Mm = TT(Mcores).matricize()
    # generalised
    lam, y, it = revp.als(TT(H), y0t, operator_gevp=TT(Mcores), repeats=6, solver='eig', sigma=-100.0)
    s['gevp_val'] = np.array(float(np.real(lam)))
    s['gevp_iters'] = np.array(it)
    _save_tt(s, 'gevp_vec', y.cores)
    s['gevp_dense'] = np.array(sla.eigh(Hm, Mm, eigvals_only=True)[0])
    # deflation of the ground state
    lam0, y_gs, _ = revp.als(TT(H), y0t, repeats=6, solver='eig', sigma=-100.0)
    lam1, y_ex, it = revp.als(TT(H), y1, previous=[y_gs], shift=50.0, repeats=6, solver='eig', sigma=-100.0)
    _save_tt(s, 'defl_prev', y_gs.cores)
    s['defl_prev_val'] = np.array(float(np.real(lam0)))
    s['defl_val'] = np.array(float(np.real(lam1)))
    s['defl_iters'] = np.array(it)
    _save_tt(s, 'defl_vec', y_ex.cores)
    s['dense_lowest2'] = np.linalg.eigvalsh(Hm)[:2].copy()
    # two eigenpairs
    vals, vecs, it = revp.als(TT(H), y0t, number_ev=2, repeats=6, solver='eig', sigma=-100.0)
    s['two_vals'] = np.real(np.array(vals, dtype=complex))
    s['two_iters'] = np.array(it)
    for e in range(2):
        _save_tt(s, 'two_vec%d' % e, vecs[e].cores)
    np.savez_compressed(os.path.join(HERE, 'evp_options.npz'), **s)
    print('evp options: ulam eig', s['ulam_ns_val'], s['ulam_ns3_vals'], s['ulam_ns_dense_top3'],
          'imag', s['ulam_ns_complex_parts'], 'ulam eigh', s['ulam_vals'], s['ulam_dense_top3'], 'gevp', s['gevp_val'], s['gevp_dense'],
          'deflation', s['defl_val'], s['dense_lowest2'], 'two', s['two_vals'])


CASES = {'base': main, 'c2': c2_instance, 'ode': ode_cases, 'power': power_case, 'mals': mals_case,
         'evp_options': evp_options_case}

if __name__ == '__main__':
    # usage: make_golden.py [reference root] [case ...]   (cases: base, c2, ode, power, mals, evp_options; default: all)
    root = sys.argv[1] if len(sys.argv) > 1 else '/root/reference'
    for name in (sys.argv[2:] or list(CASES)):
        CASES[name](root)
