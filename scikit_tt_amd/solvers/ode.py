"""Implicit integrators for linear ODEs dx/dt = A x in TT format on the device.

Drop-in for scikit_tt.solvers.ode.implicit_euler / trapezoidal_rule /
adaptive_step_size and their error functions (scikit_tt/solvers/ode.py:
249-330, 333-364, 366-450, 453-484, 487-636): each time step is an
``sle.als`` solve of (I - h A) x_{k+1} = x_k (implicit Euler) or
(I - h/2 A) x_{k+1} = (I + h/2 A) x_k (trapezoidal rule), warm-started from
the previous step's solution, with the same normalisation and step-size
rules.  The TT arithmetic (I - h A, products, differences, norms) and every
solve run on the device; ``progress`` is accepted and ignored (console
output).  ``tt_solver='mals'`` runs ``sle.mals``.
"""
import numpy as np

from .. import tensor_train as tt
from . import sle

__all__ = ['implicit_euler', 'errors_impl_euler', 'trapezoidal_rule', 'errors_trapezoidal', 'adaptive_step_size']


def _solve(tt_solver, A, x0, b, micro_solver, repeats, threshold, max_rank):
    if tt_solver == 'als':
        return sle.als(A, x0, b, solver=micro_solver, repeats=repeats)
    if tt_solver == 'mals':
        return sle.mals(A, x0, b, solver=micro_solver, repeats=repeats, threshold=threshold, max_rank=max_rank)
    raise NotImplementedError("tt_solver must be 'als' or 'mals', got %r" % (tt_solver,))


def implicit_euler(operator, initial_value, initial_guess, step_sizes, repeats=1, tt_solver='als', threshold=1e-12,
                   max_rank=np.inf, micro_solver='solve', normalize=1, progress=True):
    """Implicit Euler method (scikit_tt/solvers/ode.py:249-330)."""
    solution = [initial_value]
    tt_tmp = initial_guess
    eye = tt.eye(operator.row_dims)
    for h in step_sizes:
        tt_tmp = _solve(tt_solver, eye - h * operator, tt_tmp, solution[-1], micro_solver, repeats, threshold,
                        max_rank)
        if normalize > 0:
            tt_tmp = (1 / tt_tmp.norm(p=normalize)) * tt_tmp
        solution.append(tt_tmp.copy())
    return solution


def errors_impl_euler(operator, solution, step_sizes):
    """Relative residuals of the implicit Euler steps (scikit_tt/solvers/ode.py:333-364)."""
    eye = tt.eye(operator.row_dims)
    return [((eye - step_sizes[i] * operator).dot(solution[i + 1]) - solution[i]).norm() / solution[i].norm()
            for i in range(len(solution) - 1)]


def trapezoidal_rule(operator, initial_value, initial_guess, step_sizes, repeats=1, tt_solver='als', threshold=1e-12,
                     max_rank=np.inf, micro_solver='solve', normalize=1, progress=True):
    """Trapezoidal rule (scikit_tt/solvers/ode.py:366-450)."""
    solution = [initial_value]
    tt_tmp = initial_guess
    eye = tt.eye(operator.row_dims)
    for h in step_sizes:
        rhs = (eye + 0.5 * h * operator).dot(solution[-1])
        tt_tmp = _solve(tt_solver, eye - 0.5 * h * operator, tt_tmp, rhs, micro_solver, repeats, threshold, max_rank)
        if normalize > 0:
            tt_tmp = (1 / tt_tmp.norm(p=normalize)) * tt_tmp
        solution.append(tt_tmp.copy())
    return solution


def errors_trapezoidal(operator, solution, step_sizes):
    """Relative residuals of the trapezoidal steps (scikit_tt/solvers/ode.py:453-484)."""
    eye = tt.eye(operator.row_dims)
    errs = []
    for i in range(len(solution) - 1):
        h = step_sizes[i]
        rhs = (eye + 0.5 * h * operator).dot(solution[i])
        errs.append(((eye - 0.5 * h * operator).dot(solution[i + 1]) - rhs).norm() / rhs.norm())
    return errs


def adaptive_step_size(operator, initial_value, initial_guess, time_end, step_size_first=1e-10, repeats=1,
                       solver='solve', error_tol=1e-1, closeness_tol=0.5, step_size_min=1e-14, step_size_max=10,
                       closeness_min=1e-3, factor_max=2, factor_safe=0.9, second_method='two_step_Euler',
                       normalize=1, progress=True):
    """Adaptive step size method (scikit_tt/solvers/ode.py:487-636): an
    implicit Euler step and a second-order step (two half Euler steps or the
    trapezoidal rule) of the same size; the step is accepted when the local
    error and the relative change of ||A x|| stay within their tolerances,
    and the next size follows the same safety-factor rule."""
    solution = [initial_value]
    closeness_pre = operator.dot(initial_value).norm()
    t_tmp = initial_guess
    time = 0
    time_steps = [0]
    step_size = step_size_first
    eye = tt.eye(operator.row_dims)
    t_2 = None
    while time < time_end and closeness_pre > closeness_min and step_size > step_size_min:
        t_1 = sle.als(eye - step_size * operator, t_tmp.copy(), solution[-1], solver=solver, repeats=repeats)
        t_1 = (1 / t_1.norm(p=1)) * t_1
        if second_method == 'two_step_Euler':
            t_2 = sle.als(eye - 0.5 * step_size * operator, t_tmp.copy(), solution[-1], solver=solver,
                          repeats=repeats)
            t_2 = sle.als(eye - 0.5 * step_size * operator, t_2.copy(), solution[-1], solver=solver, repeats=repeats)
        if second_method == 'trapezoidal_rule':
            t_2 = sle.als(eye - 0.5 * step_size * operator, t_tmp.copy(),
                          (eye + 0.5 * step_size * operator).dot(solution[-1]), solver=solver, repeats=repeats)
        if normalize > 0:
            t_2 = (1 / t_2.norm(p=normalize)) * t_2
        closeness = operator.dot(t_1).norm()
        local_error = (t_1 - t_2).norm() / t_1.norm()
        closeness_difference = (closeness - closeness_pre) / closeness_pre
        factor_local = error_tol / local_error
        factor_closeness = closeness_tol / np.abs(closeness_difference)
        step_size_new = np.amin([factor_max, factor_safe * factor_local, factor_safe * factor_closeness]) * step_size
        if factor_local > 1 and factor_closeness > 1:
            time = np.min([time + step_size, time_end])
            step_size = np.amin([step_size_new, time_end - time, step_size_max])
            solution.append(t_2.copy())
            time_steps.append(time)
            t_tmp = t_1
            closeness_pre = closeness
        else:
            step_size = step_size_new
    return solution, time_steps
