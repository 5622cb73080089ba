// fe_quadrature.hip -- finite-element point source for the spread: the force
// density and the positions at the quadrature points of every element, as
// FEDataManager::spread computes them before calling LEInteractor::spread
// (ibtk/src/lagrangian/FEDataManager.cpp:1169-1204: sum_weighted_elem_solution
// :825-857, then zeroExteriorValues :2738-2783 per patch).
//
// One thread per (element, quadrature point): X_qp = sum_k X_node[k] phi[k][qp]
// and F_JxW_qp = (sum_k F_node[k] phi[k][qp]) * JxW[qp], each sum running over
// the element's basis functions in order from 0.0, as the reference loop does
// (so FP64 results are bitwise those of the reference arithmetic).  The shape
// function table (one per quadrature rule) and JxW (one per element and point)
// come from the caller's finite-element library, as they come from libMesh in
// the reference.  The quadrature points then go through a level plan like any
// other Lagrangian points: the level spread adds every point once, into the
// patch that owns its cell, which is what the per-patch spread after
// zeroExteriorValues followed by the ghost accumulation amounts to.
#include "ibamr_amd.h"

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

namespace ibamr
{
int set_error(int code, const std::string& msg);
}

namespace
{
constexpr int FE_MAX_VARS = 8;

template <typename TF>
__global__ void k_fe_qp(int ndim, int nvars, int nelem, int nbasis, int nqp, int nnodes, const int* __restrict__ elem_nodes,
                        const double* __restrict__ phi, const double* __restrict__ JxW,
                        const double* __restrict__ X_nodes, const double* __restrict__ F_nodes,
                        double* __restrict__ X_qp, TF* __restrict__ F_qp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (long)nelem * nqp) return;
    const int e = (int)(t / nqp), q = (int)(t - (long)e * nqp);
    double x[3] = {0.0, 0.0, 0.0};
    double f[FE_MAX_VARS];
#pragma unroll
    for (int i = 0; i < FE_MAX_VARS; ++i) f[i] = 0.0;
    const int* en = elem_nodes + (size_t)e * nbasis;
    bool bad = false;
    for (int k = 0; k < nbasis; ++k)
    {
        const int nk = en[k];
        if (nk < 0 || nk >= nnodes)  // never read outside the node arrays; the outputs become NaN
        {
            bad = true;
            continue;
        }
        const size_t node = (size_t)nk;
        const double p = phi[(size_t)k * nqp + q];
        for (int d = 0; d < ndim; ++d) x[d] = x[d] + X_nodes[node * ndim + d] * p;
        if (F_nodes)
#pragma unroll
            for (int i = 0; i < FE_MAX_VARS; ++i)
                if (i < nvars) f[i] = f[i] + F_nodes[node * nvars + i] * p;
    }
    if (bad)
    {
        x[0] = x[1] = x[2] = __builtin_nan("");
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i) f[i] = __builtin_nan("");
    }
    for (int d = 0; d < ndim; ++d) X_qp[(size_t)t * ndim + d] = x[d];
    if (F_nodes)
    {
        const double w = JxW ? JxW[(size_t)e * nqp + q] : 1.0;
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i)
            if (i < nvars) F_qp[(size_t)t * nvars + i] = (TF)(JxW ? f[i] * w : f[i]);
    }
}
// FEDataManager::interpWeighted after the interpolation to the quadrature
// points (FEDataManager.cpp:2009-2061, integrate_elem_rhs :1569-1606): every
// element forms rhs[k][i] = sum_q U_qp[q][i] * (phi[k][q] * JxW[q]) (q in order
// from 0.0) and adds it to its nodes in element order.  One thread per node
// walks its (element, basis) list in element order, so each node value is that
// same sequence of operations.
template <typename TU>
__global__ void k_fe_node_rhs(int nvars, int nnodes, int nqp, const int* __restrict__ start,
                              const int2* __restrict__ ent, const double* __restrict__ phi,
                              const double* __restrict__ JxW, const TU* __restrict__ U_qp, double* __restrict__ F,
                              int overwrite)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nnodes) return;
    double acc[FE_MAX_VARS];
#pragma unroll
    for (int i = 0; i < FE_MAX_VARS; ++i) acc[i] = overwrite ? 0.0 : (i < nvars ? F[(size_t)v * nvars + i] : 0.0);
    for (int j = start[v]; j < start[v + 1]; ++j)
    {
        const int e = ent[j].x, k = ent[j].y;
        double r[FE_MAX_VARS];
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i) r[i] = 0.0;
        for (int q = 0; q < nqp; ++q)
        {
            const double p = phi[(size_t)k * nqp + q] * JxW[(size_t)e * nqp + q];
            const TU* u = U_qp + ((size_t)e * nqp + q) * nvars;
#pragma unroll
            for (int i = 0; i < FE_MAX_VARS; ++i)
                if (i < nvars) r[i] = r[i] + (double)u[i] * p;
        }
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i) acc[i] = acc[i] + r[i];
    }
#pragma unroll
    for (int i = 0; i < FE_MAX_VARS; ++i)
        if (i < nvars) F[(size_t)v * nvars + i] = acc[i];
}
} // namespace

struct ibamr_fe_mesh
{
    int nelem = 0, nbasis = 0, nnodes = 0;
    int* start = nullptr;   // [nnodes+1]
    int2* ent = nullptr;    // (element, basis) per node, element order
};

extern "C" {

int ibamr_fe_mesh_destroy(ibamr_fe_mesh* m)
{
    if (!m) return IBAMR_OK;
    if (m->start) (void)hipFree(m->start);
    if (m->ent) (void)hipFree(m->ent);
    delete m;
    return IBAMR_OK;
}

int ibamr_fe_mesh_create(ibamr_fe_mesh** out, int nelem, int nbasis, int nnodes, const int* elem_nodes)
{
    using ibamr::set_error;
    if (!out) return set_error(IBAMR_ERR_INVALID, "null argument");
    *out = nullptr;
    if (nelem < 0 || nbasis < 1 || nnodes < 0) return set_error(IBAMR_ERR_INVALID, "bad mesh sizes");
    if (nelem > 0 && !elem_nodes) return set_error(IBAMR_ERR_INVALID, "null connectivity");
    std::vector<int> cnt((size_t)nnodes + 1, 0);
    for (long t = 0; t < (long)nelem * nbasis; ++t)
    {
        const int v = elem_nodes[t];
        if (v < 0 || v >= nnodes)
            return set_error(IBAMR_ERR_INVALID, "element " + std::to_string(t / nbasis) + ": bad node index");
        ++cnt[(size_t)v + 1];
    }
    for (int v = 0; v < nnodes; ++v) cnt[(size_t)v + 1] += cnt[v];
    std::vector<int2> ent((size_t)std::max(1L, (long)nelem * nbasis));
    std::vector<int> pos(cnt.begin(), cnt.end() - 1);
    for (int e = 0; e < nelem; ++e)
        for (int k = 0; k < nbasis; ++k) ent[(size_t)pos[elem_nodes[(size_t)e * nbasis + k]]++] = make_int2(e, k);
    ibamr_fe_mesh* m = new ibamr_fe_mesh();
    m->nelem = nelem;
    m->nbasis = nbasis;
    m->nnodes = nnodes;
    if (hipMalloc(&m->start, sizeof(int) * cnt.size()) != hipSuccess ||
        hipMalloc(&m->ent, sizeof(int2) * ent.size()) != hipSuccess ||
        hipMemcpy(m->start, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(m->ent, ent.data(), sizeof(int2) * ent.size(), hipMemcpyHostToDevice) != hipSuccess)
    {
        ibamr_fe_mesh_destroy(m);
        return set_error(IBAMR_ERR_DEVICE, "device allocation or copy failed for the FE mesh");
    }
    *out = m;
    return IBAMR_OK;
}

int ibamr_fe_integrate_rhs(const ibamr_fe_mesh* m, int nvars, int nqp, const double* phi, const double* JxW,
                           const void* U_qp, int u_precision, double* F_nodes, int mode, void* stream)
{
    using ibamr::set_error;
    if (!m) return set_error(IBAMR_ERR_INVALID, "null mesh");
    if (nvars < 1 || nvars > FE_MAX_VARS) return set_error(IBAMR_ERR_INVALID, "nvars must be 1..8");
    if (nqp < 1) return set_error(IBAMR_ERR_INVALID, "nqp < 1");
    if (mode != IBAMR_SPREAD_ACCUMULATE && mode != IBAMR_SPREAD_OVERWRITE) return set_error(IBAMR_ERR_INVALID, "bad mode");
    if (u_precision != IBAMR_FP64 && u_precision != IBAMR_FP32C) return set_error(IBAMR_ERR_INVALID, "bad precision");
    if (m->nnodes == 0) return IBAMR_OK;
    if (!F_nodes || (m->nelem > 0 && (!phi || !JxW || !U_qp))) return set_error(IBAMR_ERR_INVALID, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256;
    const dim3 G((m->nnodes + B - 1) / B);
    const int ow = mode == IBAMR_SPREAD_OVERWRITE ? 1 : 0;
    if (u_precision == IBAMR_FP64)
        hipLaunchKernelGGL(k_fe_node_rhs<double>, G, dim3(B), 0, s, nvars, m->nnodes, nqp, m->start, m->ent, phi, JxW,
                           static_cast<const double*>(U_qp), F_nodes, ow);
    else
        hipLaunchKernelGGL(k_fe_node_rhs<float>, G, dim3(B), 0, s, nvars, m->nnodes, nqp, m->start, m->ent, phi, JxW,
                           static_cast<const float*>(U_qp), F_nodes, ow);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_node_rhs: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

int ibamr_fe_quadrature_points(int ndim, int nvars, int nelem, int nbasis, int nqp, int nnodes, const int* elem_nodes,
                               const double* phi, const double* JxW, const double* X_nodes, const double* F_nodes,
                               double* X_qp, void* F_qp, int f_precision, void* stream)
{
    using ibamr::set_error;
    if (ndim != 2 && ndim != 3) return set_error(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    if (nelem < 0 || nbasis < 1 || nqp < 1 || nnodes < 0) return set_error(IBAMR_ERR_INVALID, "bad element sizes");
    if (F_nodes && (nvars < 1 || nvars > FE_MAX_VARS))
        return set_error(IBAMR_ERR_INVALID, "nvars must be 1.." + std::to_string(FE_MAX_VARS));
    if ((long)nelem * nqp > 0x7fffffffL) return set_error(IBAMR_ERR_CAPACITY, "too many quadrature points");
    if (nelem == 0) return IBAMR_OK;
    if (!elem_nodes || !phi || !X_nodes || !X_qp || (F_nodes && !F_qp))
        return set_error(IBAMR_ERR_INVALID, "null argument");
    if (f_precision != IBAMR_FP64 && f_precision != IBAMR_FP32C) return set_error(IBAMR_ERR_INVALID, "bad precision");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const long tot = (long)nelem * nqp;
    const int B = 256;
    const dim3 G((unsigned)((tot + B - 1) / B));
    if (f_precision == IBAMR_FP64)
        hipLaunchKernelGGL(k_fe_qp<double>, G, dim3(B), 0, s, ndim, nvars, nelem, nbasis, nqp, nnodes, elem_nodes, phi,
                           JxW, X_nodes, F_nodes, X_qp, static_cast<double*>(F_qp));
    else
        hipLaunchKernelGGL(k_fe_qp<float>, G, dim3(B), 0, s, ndim, nvars, nelem, nbasis, nqp, nnodes, elem_nodes, phi,
                           JxW, X_nodes, F_nodes, X_qp, static_cast<float*>(F_qp));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_qp: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

} // extern "C"
