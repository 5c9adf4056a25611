/*
 * Right-hand side and exact Jacobian of the chemical network.
 * Test infrastructure (see pdr_oracle.h).
 *
 * The RHS is the MakeRates form of odes_*.c:26-148 evaluated from the network
 * tables: flux_r = k_r * X[a_r] * X[b_r] * (n_H for two-body and grain
 * reactions), ydot_i = sum_r S_ir flux_r with net stoichiometry S, and
 * x_e = sum_i charge_i x_i (odes_reduced.c:44).  The reference writes each
 * equation as form + x_i*loss without cancelling catalysts; the sums are the
 * same in exact arithmetic.
 *
 * The Jacobian is exact, including d x_e / d x_j = charge_j, which the
 * shipped jacobian_*.c omits (jacobian_reduced.c:41 freezes x_e).
 */
#include <string.h>

#include "pdr_oracle.h"

static inline double xval(const pdr_net *n, const double *x, double xe, int k) {
  if (k < n->neq) return x[k];
  return (k == n->neq) ? xe : 1.0;
}

void pdr_rhs(const pdr_net *n, const double *rate, double n_H, const double *x, double *ydot,
             double *x_e) {
  const int neq = n->neq;
  double xe = 0.0;
  for (int i = 0; i < neq; i++)
    if (n->charge[i]) xe += n->charge[i] * x[i];
  for (int i = 0; i < neq; i++) ydot[i] = 0.0;
  for (int r = 0; r < n->nreac; r++) {
    int s0 = n->st_ptr[r], s1 = n->st_ptr[r + 1];
    if (s0 == s1) continue;
    double k = rate[r];
    if (n->nh[r]) k *= n_H;
    double fl = k * xval(n, x, xe, n->fa[r]) * xval(n, x, xe, n->fb[r]);
    for (int s = s0; s < s1; s++) ydot[n->st_idx[s]] += n->st_coef[s] * fl;
  }
  if (x_e) *x_e = xe;
}

void pdr_jac(const pdr_net *n, const double *rate, double n_H, const double *x, double *J) {
  const int neq = n->neq;
  double xe = 0.0;
  for (int i = 0; i < neq; i++)
    if (n->charge[i]) xe += n->charge[i] * x[i];
  memset(J, 0, sizeof(double) * (size_t)neq * neq);
  /* g_i = d ydot_i / d x_e (applied as a rank-1 update with the charges) */
  double gp[1024];
  if (neq > 1024) return;
  for (int i = 0; i < neq; i++) gp[i] = 0.0;
  for (int r = 0; r < n->nreac; r++) {
    int s0 = n->st_ptr[r], s1 = n->st_ptr[r + 1];
    if (s0 == s1) continue;
    double k = rate[r];
    if (n->nh[r]) k *= n_H;
    int a = n->fa[r], b = n->fb[r];
    double xa = xval(n, x, xe, a), xb = xval(n, x, xe, b);
    /* partials of the flux with respect to its two factors */
    double da = k * xb, db = k * xa;
    for (int s = s0; s < s1; s++) {
      int i = n->st_idx[s];
      double c = n->st_coef[s];
      if (a < neq) J[(size_t)i * neq + a] += c * da;
      else if (a == neq) gp[i] += c * da;
      if (b < neq) J[(size_t)i * neq + b] += c * db;
      else if (b == neq) gp[i] += c * db;
    }
  }
  for (int i = 0; i < neq; i++) {
    if (gp[i] == 0.0) continue;
    for (int j = 0; j < neq; j++)
      if (n->charge[j]) J[(size_t)i * neq + j] += gp[i] * n->charge[j];
  }
}

void pdr_jac_bordered(const pdr_net *n, const double *rate, double n_H, const double *x,
                      double *Jb) {
  const int neq = n->neq, nb = neq + 1;
  double xe = 0.0;
  for (int i = 0; i < neq; i++)
    if (n->charge[i]) xe += n->charge[i] * x[i];
  memset(Jb, 0, sizeof(double) * (size_t)nb * nb);
  for (int r = 0; r < n->nreac; r++) {
    int s0 = n->st_ptr[r], s1 = n->st_ptr[r + 1];
    if (s0 == s1) continue;
    double k = rate[r];
    if (n->nh[r]) k *= n_H;
    int a = n->fa[r], b = n->fb[r];
    double xa = xval(n, x, xe, a), xb = xval(n, x, xe, b);
    double da = k * xb, db = k * xa;
    for (int s = s0; s < s1; s++) {
      int i = n->st_idx[s];
      double c = n->st_coef[s];
      if (a <= neq) Jb[(size_t)i * nb + a] += c * da;  /* column neq is x_e */
      if (b <= neq) Jb[(size_t)i * nb + b] += c * db;
    }
  }
}
