// Single reference update kernels on the device (FP64), for callers that
// drive BaseFDTD11's per-step functions one at a time instead of an
// integrator.  Each op is one elementwise pass over index ranges, written in
// the reference's expression order with FP contraction off, so the result
// is the reference's bit for bit:
//   OP_EX_UPDATE   ADE_ExUpdate        BaseFDTD11.py:663-669
//   OP_HY_UPDATE   ADE_HyUpdate        :640-656
//   OP_PSI_E       CPML_Psi_e_Update   :364-376
//   OP_PSI_M       CPML_Psi_m_Update   :381-393
//   OP_DX_UPDATE   ADE_DxUpdate        :750-760
//   OP_EX_CREATE   ADE_ExCreate        :712-725
//   OP_POLARISE    ADE_PolarisationCurrent_Ex :609-633
//   OP_ACUBIC      AcubicFinder + Nonlin_Cubic_Solver :803-853 (root 0, 1728.org form)
//   OP_NONLIN_EX   NonLinExUpdate      :858-877
// and the per-step functions of BaseFDTD.py (the older module: other index
// ranges, the psi_E correction added, the psi_H difference reversed):
//   OP_B_EX        CPML_ExUpdate :355-358, ADE_ExUpdate :401-408
//   OP_B_PSI_E     CPML_Psi_e_Update   :331-336
//   OP_B_PSI_M     CPML_Psi_m_Update   :339-344
//   OP_B_POLARISE  ADE_PolarisationCurrent_Ex :378-390
//   OP_B_DX        ADE_DxUpdate        :417-422
// (BaseFDTD's CPML_HyUpdate and ADE_HyUpdate are OP_HY_UPDATE's form, and its
// ADE_ExCreate is OP_EX_CREATE's.)
#include <hip/hip_runtime.h>

#include "device_common.hpp"

namespace fdtd {

struct OpArgs {
    double* a[8];
    double s[4];
    int lo[2], hi[2];       // up to two index ranges (the CPML strips); hi <= lo: empty
};

__global__ __launch_bounds__(256) void step_op(int op, OpArgs g) {
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const int n0 = g.hi[0] > g.lo[0] ? g.hi[0] - g.lo[0] : 0;
    const int n1 = g.hi[1] > g.lo[1] ? g.hi[1] - g.lo[1] : 0;
    if (t >= n0 + n1) return;
    const int n = t < n0 ? g.lo[0] + t : g.lo[1] + (t - n0);
    double** a = g.a;
    switch (op) {
        case 0:   // Ex[n] += (Hy[n] - Hy[n-1] - Jx[n]) * UpExMat[n] * den_Exdz[n]
            a[0][n] = a[0][n] + (a[1][n] - a[1][n - 1] - a[2][n]) * a[3][n] * a[4][n];
            break;
        case 1:   // Hy[n] = Hy[n]*UpHySelf[n] + (Ex[n+1] - Ex[n]) * UpHyMat[n] * den_Hydz[n]
            a[0][n] = a[0][n] * a[2][n] + (a[1][n + 1] - a[1][n]) * a[3][n] * a[4][n];
            break;
        case 2:   // psi_Ex = beX psi_Ex + ceX (Hy[n] - Hy[n-1]); Ex -= Cb psi_Ex
            a[1][n] = a[3][n] * a[1][n] + a[4][n] * (a[2][n] - a[2][n - 1]);
            a[0][n] = a[0][n] - a[5][n] * a[1][n];
            break;
        case 3:   // psi_Hy = bmY psi_Hy + cmY (Ex[n+1] - Ex[n]); Hy += C2 psi_Hy
            a[1][n] = a[3][n] * a[1][n] + a[4][n] * (a[2][n + 1] - a[2][n]);
            a[0][n] = a[0][n] + a[5][n] * a[1][n];
            break;
        case 4:   // Dx[n] += (Hy[n] - Hy[n-1]) * (dt/dz) * den_Exdz[n]
            a[0][n] = a[0][n] + (a[1][n] - a[1][n - 1]) * g.s[0] * a[2][n];
            break;
        case 5:   // Ex[n] = (Dx[n] - P[n]) / eps0
            a[0][n] = (a[1][n] - a[2][n]) / g.s[0];
            break;
        case 6:   // P[n] = A P[n] + B P_prev[n] + C Ex[n]
            a[0][n] = g.s[0] * a[0][n] + g.s[1] * a[1][n] + g.s[2] * a[2][n];
            break;
        case 7: {  // Acubic[n] = root 0 of (cub, qua, one, -|Dx/eps0|^2)
            const double dd = a[1][n] / g.s[3];
            a[0][n] = cubic_root0_1728(g.s[0], g.s[1], g.s[2], -(fabs(dd) * fabs(dd)));
            break;
        }
        case 8:   // Ex[n] = Dx[n] / (eps0 * sqrt(1.2) + eps0 * chi3 * Acubic[n])
            a[0][n] = a[1][n] / (g.s[0] * g.s[2] + g.s[0] * g.s[1] * a[2][n]);
            break;
        case 9:   // Ex[n] = Ex[n]*Self[n] + (Hy[n] - Hy[n-1]) * c1[n] * c2[n]
            a[0][n] = a[0][n] * a[2][n] + (a[1][n] - a[1][n - 1]) * a[3][n] * a[4][n];
            break;
        case 10:  // psi_Ex = beX psi_Ex + ceX (Hy[n] - Hy[n-1]); Ex += Cb psi_Ex
            a[1][n] = a[3][n] * a[1][n] + a[4][n] * (a[2][n] - a[2][n - 1]);
            a[0][n] = a[0][n] + a[5][n] * a[1][n];
            break;
        case 11:  // psi_Hy = bmY psi_Hy + cmY (Ex[n] - Ex[n+1]); Hy += C2 psi_Hy
            a[1][n] = a[3][n] * a[1][n] + a[4][n] * (a[2][n] - a[2][n + 1]);
            a[0][n] = a[0][n] + a[5][n] * a[1][n];
            break;
        case 12:  // P[n] = A Pprev[n] + B Pprev2[n] + C Ex[n]
            a[0][n] = g.s[0] * a[1][n] + g.s[1] * a[2][n] + g.s[2] * a[3][n];
            break;
        case 13:  // Dx[n] += (Hy[n] - Hy[n-1]) * dt / dz; Ex[n] = (Dx[n] - P[n]) / eps0
            a[0][n] = a[0][n] + (a[2][n] - a[2][n - 1]) * g.s[0] / g.s[1];
            a[1][n] = (a[0][n] - a[3][n]) / g.s[2];
            break;
        default:
            break;
    }
}

hipError_t launch_step_op(int op, const OpArgs& g, hipStream_t st) {
    const int n0 = g.hi[0] > g.lo[0] ? g.hi[0] - g.lo[0] : 0;
    const int n1 = g.hi[1] > g.lo[1] ? g.hi[1] - g.lo[1] : 0;
    const int n = n0 + n1;
    if (n == 0) return hipSuccess;
    hipLaunchKernelGGL(step_op, dim3((n + 255) / 256), dim3(256), 0, st, op, g);
    return hipGetLastError();
}

}  // namespace fdtd
