"""torch.autograd binding of the converged Newton solution (beside codes/PythonTools).

`solved_observables(system, ome)` solves F(u, Ome) = 0 on the device (do_newton
until ||F||_inf < precision) and returns Q = (M_ADM, M_Komar, J_ADM) as a float64
tensor.  Its backward is the adjoint method of the implicit solution:
J^T lambda_q = dQ_q/du with the LU factors of J (kad_xcts_sensitivity), and
dQ_q/dOme = -lambda_q . dF/dOme; nothing is differentiated through the Newton
iterations, and no operator intermediate is stored.
"""
import torch


class _SolvedObservables(torch.autograd.Function):
    @staticmethod
    def forward(ctx, ome, system, precision, solver, guess):
        system.set_ome(float(ome))
        system.solve(precision, solver, guess)
        q = system.observables()
        sens = system.sensitivity()
        ctx.save_for_backward(torch.tensor([sens["M_adm"], sens["M_komar"], sens["J_adm"]], dtype=torch.float64))
        return torch.tensor([q["M_adm"], q["M_komar"], q["J_adm"]], dtype=torch.float64)

    @staticmethod
    def backward(ctx, grad_q):
        (dq,) = ctx.saved_tensors
        return (grad_q.to(torch.float64) * dq).sum(), None, None, None, None


def solved_observables(system, ome, precision=1e-10, solver="f64", guess=None):
    """Q(Ome) = (M_ADM, M_Komar, J_ADM) at the solution for horizon rotation `ome`
    (a 0-d float64 tensor; requires_grad gives dQ/dOme by the adjoint solve)."""
    return _SolvedObservables.apply(ome, system, precision, solver, guess)
