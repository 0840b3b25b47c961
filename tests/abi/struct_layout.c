/* Prints the size and field offsets of the v2g.h structures, one
 * "name.field offset" line each, so the ctypes mirror (v2rdm_casscf_amd/_lib.py)
 * can be checked against what a C compiler makes of the header. */
#include <stddef.h>
#include <stdio.h>

#include "../../include/v2g.h"

#define F(T, f) printf(#T "." #f " %zu\n", offsetof(T, f))
#define S(T) printf(#T ".sizeof %zu\n", sizeof(T))

int main(void)
{
    S(v2g_layout_desc);
    F(v2g_layout_desc, nirrep); F(v2g_layout_desc, amopi); F(v2g_layout_desc, nalpha);
    F(v2g_layout_desc, nbeta); F(v2g_layout_desc, multiplicity);
    F(v2g_layout_desc, constrain_q2); F(v2g_layout_desc, constrain_g2);
    F(v2g_layout_desc, constrain_t2); F(v2g_layout_desc, constrain_spin);
    F(v2g_layout_desc, constrain_sz); F(v2g_layout_desc, constrain_t1);
    F(v2g_layout_desc, spin_adapt_q2); F(v2g_layout_desc, spin_adapt_g2);
    F(v2g_layout_desc, energy_order);
    S(v2g_params);
    F(v2g_params, tau); F(v2g_params, r_convergence); F(v2g_params, e_convergence);
    F(v2g_params, cg_convergence); F(v2g_params, maxiter); F(v2g_params, cg_maxiter);
    F(v2g_params, mu_update_frequency); F(v2g_params, precond); F(v2g_params, flexible);
    F(v2g_params, verbose); F(v2g_params, max_seconds); F(v2g_params, train_iters);
    F(v2g_params, train_batch); F(v2g_params, train_steps); F(v2g_params, train_lr);
    F(v2g_params, eig_warm_sweeps); F(v2g_params, eig_inner_sweeps);
    F(v2g_params, cg_extrapolate); F(v2g_params, mu0); F(v2g_params, oiter0);
    F(v2g_params, ep0); F(v2g_params, ed0); F(v2g_params, mlp_base);
    F(v2g_params, train_penalty);
    F(v2g_params, eig_pair_order);
    S(v2g_stats);
    F(v2g_stats, oiter); F(v2g_stats, cg_iters); F(v2g_stats, converged);
    F(v2g_stats, energy_primal); F(v2g_stats, energy_dual); F(v2g_stats, ep); F(v2g_stats, ed);
    F(v2g_stats, mu); F(v2g_stats, t_cg); F(v2g_stats, t_xz); F(v2g_stats, t_total);
    F(v2g_stats, eig_sweeps); F(v2g_stats, precond_active); F(v2g_stats, cg_iters_train);
    F(v2g_stats, train_outer); F(v2g_stats, mlp_loss); F(v2g_stats, fallback);
    F(v2g_stats, exact_splits);
    return 0;
}
