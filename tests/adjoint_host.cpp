// Host build of the generated per-cell primal/VJP routines, for the CPU tests.
#include "hes_adjoint_gen.hpp"
extern "C" {
void gen_cell_fwd(const double* in, double* out) { hes_gen::cell_fwd(in, out); }
void gen_cell_vjp(const double* in, const double* ob, double* ib) { hes_gen::cell_vjp(in, ob, ib); }
// the fused primal + VJP with fixed seeds; tail = whether the acoustic and u1
// seeds are used (the acoustic tensor's primal runs only then)
void gen_cell_fused(const double* in, double* out, const double* ob, int tail, double* ib) {
  hes_gen::cell_fwd_vjp(in, out,
      [&](const double*, double* o) {
        for (int k = 0; k < hes_gen::CELL_OUT; ++k) o[k] = (k < 14 || tail) ? ob[k] : 0.0;
        return tail != 0;
      },
      [&](const double*, double* o) {
        for (int k = 14; k < 20; ++k) o[k] = ob[k];
      },
      ib);
}
void gen_p2c_fwd(const double* in, double* out) { hes_gen::prim2cons_fwd(in, out); }
void gen_p2c_vjp(const double* in, const double* ob, double* ib) { hes_gen::prim2cons_vjp(in, ob, ib); }
void gen_loss_fwd(const double* in, double* out) { hes_gen::loss_fwd(in, out); }
void gen_loss_vjp(const double* in, const double* ob, double* ib) { hes_gen::loss_vjp(in, ob, ib); }
}
