// Solver assemblies (solvers.hip); all pointers are device memory.
#pragma once
#include "osc_internal.h"

namespace osc {
void solver2(osc_ctx* c, const int n[3], const double* rho, double gamma, const double delta[3],
             double* ef, double* phi, const double* offset);
void space_charge_3d(osc_ctx* c, const osc_mesh& m, const double* rho, const double* offset,
                     int at_cathode, int calc_b, double* ef, double* bf);
int legacy_free(osc_ctx* c, const osc_mesh& m, const double* rho, int direct, int igf,
                double* phi, double* ef, double* bf);
int cathode_legacy(osc_ctx* c, const osc_mesh& m, double* rho, int direct, int igf, int method,
                   double* phi, double* ef, double* bf);
int rectpipe(osc_ctx* c, const osc_mesh& m, int direct, const double* rho, double* phi, double* ef,
             double* bf);
void pipe_build(osc_ctx* c, double gamma, double a, double b, const double delta[3],
                const double umin[3], const int npad[3]);
extern const int kPipeSigns[4][3];
}  // namespace osc
