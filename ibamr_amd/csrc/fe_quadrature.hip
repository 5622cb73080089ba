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
//
// Also here: a rule per element (the reference keys every element's rule
// separately, FEDataManager.cpp:1153-1165, 1951-1963, and with adaptive
// quadrature the order follows the deformed element's longest edge,
// getQuadratureKey, ibtk/src/utilities/libmesh_utilities.cpp:187-225), computed
// per element on the GPU; and nodal quadrature (:996-1075, 1761-1905): the
// nodes are the point source, weighted by the diagonal mass of the nodal rule
// (FEProjector::buildDiagonalL2MassMatrix, FEProjector.cpp:587-672).
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

// One quadrature point: x = sum_k X_node[k] ph[k*nq + q], f = sum_k F_node[k] ph[k*nq + q]
// (basis order, from 0.0); false when a node index is outside [0, nnodes).
__device__ inline bool fe_point(int ndim, int nvars, int nbasis, int nnodes, const int* __restrict__ en,
                                const double* __restrict__ ph, int nq, int q, const double* __restrict__ X_nodes,
                                const double* __restrict__ F_nodes, double* x, double* f)
{
    bool ok = true;
    for (int k = 0; k < nbasis; ++k)
    {
        const int nk = en[k];
        if (nk < 0 || nk >= nnodes)
        {
            ok = false;
            continue;
        }
        const size_t node = (size_t)nk;
        const double p = ph[(size_t)k * nq + q];
        for (int d = 0; d < ndim; ++d) x[d] = x[d] + X_nodes[node * ndim + d] * p;
        if (F_nodes)
#pragma unroll
            for (int i = 0; i < FE_MAX_VARS; ++i)
                if (i < nvars) f[i] = f[i] + F_nodes[node * nvars + i] * p;
    }
    return ok;
}

// The point source with a rule per element: point t belongs to the element e
// with qp_off[e] <= t < qp_off[e+1] (binary search); element e's rule r =
// elem_rule[e] has rule_nqp[r] points and its table at phi + rule_phi_off[r].
// An element whose rule index, point count or node indices do not fit gives
// NaN outputs and reads nothing outside the arrays.
template <typename TF>
__global__ void k_fe_qp_rules(int ndim, int nvars, int nelem, int nbasis, int nnodes, int nrules, long total,
                              const int* __restrict__ elem_nodes, const int* __restrict__ elem_rule,
                              const int* __restrict__ rule_nqp, const long* __restrict__ rule_phi_off,
                              const double* __restrict__ phi, const long* __restrict__ qp_off,
                              const double* __restrict__ JxW, const double* __restrict__ X_nodes,
                              const double* __restrict__ F_nodes, double* __restrict__ X_qp, TF* __restrict__ F_qp)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= total) return;
    int lo = 0, hi = nelem;  // the last e with qp_off[e] <= t
    while (hi - lo > 1)
    {
        const int mid = (lo + hi) >> 1;
        if (qp_off[mid] <= t) lo = mid;
        else hi = mid;
    }
    const int e = lo;
    const int q = (int)(t - qp_off[e]);
    const int r = elem_rule[e];
    double x[3] = {0.0, 0.0, 0.0};
    double f[FE_MAX_VARS];
#pragma unroll
    for (int i = 0; i < FE_MAX_VARS; ++i) f[i] = 0.0;
    bool ok = r >= 0 && r < nrules && qp_off[e + 1] - qp_off[e] == (long)rule_nqp[r] && q < rule_nqp[r];
    if (ok)
        ok = fe_point(ndim, nvars, nbasis, nnodes, elem_nodes + (size_t)e * nbasis, phi + rule_phi_off[r],
                      rule_nqp[r], q, X_nodes, F_nodes, x, f);
    if (!ok)
    {
        x[0] = x[1] = x[2] = __builtin_nan("");
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i) f[i] = __builtin_nan("");
    }
    for (int d = 0; d < ndim; ++d) X_qp[(size_t)t * ndim + d] = x[d];
    if (F_nodes)
    {
        const double w = JxW ? JxW[t] : 1.0;
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i)
            if (i < nvars) F_qp[(size_t)t * nvars + i] = (TF)(JxW ? f[i] * w : f[i]);
    }
}

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
    // never reads outside the node arrays; a bad node index makes the outputs NaN
    const bool bad = !fe_point(ndim, nvars, nbasis, nnodes, elem_nodes + (size_t)e * nbasis, phi, nqp, q, X_nodes,
                               F_nodes, x, f);
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
// The same with a rule per element (see k_fe_qp_rules); an element whose rule
// does not fit its point range makes its nodes' sums NaN.
template <typename TU>
__global__ void k_fe_node_rhs_rules(int nvars, int nnodes, int nrules, const int* __restrict__ start,
                                    const int2* __restrict__ ent, const int* __restrict__ elem_rule,
                                    const int* __restrict__ rule_nqp, const long* __restrict__ rule_phi_off,
                                    const double* __restrict__ phi, const long* __restrict__ qp_off,
                                    const double* __restrict__ JxW, const TU* __restrict__ U_qp,
                                    double* __restrict__ F, int overwrite)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nnodes) return;
    double acc[FE_MAX_VARS];
#pragma unroll
    for (int i = 0; i < FE_MAX_VARS; ++i) acc[i] = overwrite ? 0.0 : (i < nvars ? F[(size_t)v * nvars + i] : 0.0);
    for (int j = start[v]; j < start[v + 1]; ++j)
    {
        const int e = ent[j].x, k = ent[j].y;
        const int r = elem_rule[e];
        double rr[FE_MAX_VARS];
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i) rr[i] = 0.0;
        if (r < 0 || r >= nrules || qp_off[e + 1] - qp_off[e] != (long)rule_nqp[r])
        {
#pragma unroll
            for (int i = 0; i < FE_MAX_VARS; ++i) rr[i] = __builtin_nan("");
        }
        else
        {
            const int nq = rule_nqp[r];
            const double* ph = phi + rule_phi_off[r] + (size_t)k * nq;
            const long o = qp_off[e];
            for (int q = 0; q < nq; ++q)
            {
                const double p = ph[q] * JxW[o + q];
                const TU* u = U_qp + (size_t)(o + q) * nvars;
#pragma unroll
                for (int i = 0; i < FE_MAX_VARS; ++i)
                    if (i < nvars) rr[i] = rr[i] + (double)u[i] * p;
            }
        }
#pragma unroll
        for (int i = 0; i < FE_MAX_VARS; ++i) acc[i] = acc[i] + rr[i];
    }
#pragma unroll
    for (int i = 0; i < FE_MAX_VARS; ++i)
        if (i < nvars) F[(size_t)v * nvars + i] = acc[i];
}

// Diagonal mass of a (nodal) rule: M[v] = sum over v's (element, basis k) in
// element order of sum_q (phi[k][q] * phi[k][q]) * JxW[e][q] (each from 0.0),
// as FEProjector::buildDiagonalL2MassMatrix adds the element vectors.
__global__ void k_fe_nodal_mass(int nnodes, int nqp, const int* __restrict__ start, const int2* __restrict__ ent,
                                const double* __restrict__ phi, const double* __restrict__ JxW, double* __restrict__ M)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= nnodes) return;
    double acc = 0.0;
    for (int j = start[v]; j < start[v + 1]; ++j)
    {
        const int e = ent[j].x, k = ent[j].y;
        double m = 0.0;
        for (int q = 0; q < nqp; ++q)
        {
            const double p = phi[(size_t)k * nqp + q];
            m = m + (p * p) * JxW[(size_t)e * nqp + q];
        }
        acc = acc + m;
    }
    M[v] = acc;
}

// out[n][i] (=, or +=) V[n][i] * M[n]: F (x) dX for the nodal spread
// (pointwise_mult, FEDataManager.cpp:1014) and U * dX into the right-hand side
// for the nodal interpolation (:1883-1899).
template <typename TV, typename TO>
__global__ void k_fe_nodal_weight(long n, int nvars, const TV* __restrict__ V, const double* __restrict__ M,
                                  TO* __restrict__ out, int overwrite)
{
    const long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n * nvars) return;
    const double w = (double)V[t] * M[t / nvars];
    out[t] = overwrite ? (TO)w : (TO)((double)out[t] + w);
}

// get_max_edge_length (ibtk/include/ibtk/libmesh_utilities.h:1373-1511): node
// pairs (or triples through a mid-edge node) per element type.
__device__ inline double fe_d2(int ndim, const double* __restrict__ X, int a, int b)
{
    double s = 0.0;
    for (int d = 0; d < ndim; ++d)
    {
        const double t = X[(size_t)a * ndim + d] - X[(size_t)b * ndim + d];
        s = s + t * t;
    }
    return s;
}

__constant__ int c_pairs_tet4[6][2] = {{0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};
__constant__ int c_pairs_hex[12][2] = {{0, 1}, {0, 3}, {0, 4}, {1, 2}, {1, 5}, {2, 3},
                                       {2, 6}, {3, 7}, {4, 5}, {4, 7}, {5, 6}, {6, 7}};
__constant__ int c_pairs_tri[3][2] = {{0, 1}, {1, 2}, {0, 2}};
__constant__ int c_pairs_quad[4][2] = {{0, 1}, {1, 2}, {2, 3}, {0, 3}};
__constant__ int c_trip_tet10[6][3] = {{0, 4, 1}, {0, 6, 2}, {0, 7, 3}, {1, 5, 2}, {1, 8, 3}, {2, 9, 3}};
__constant__ int c_trip_hex27[12][3] = {{0, 8, 1},  {0, 11, 3}, {0, 12, 4}, {1, 9, 2},  {1, 13, 5}, {2, 10, 3},
                                        {2, 14, 6}, {3, 15, 7}, {4, 16, 5}, {4, 19, 7}, {5, 17, 6}, {6, 18, 7}};

__device__ double fe_max_edge(int type, int ndim, const int* __restrict__ n, const double* __restrict__ X)
{
    if (type == IBAMR_FE_TET10 || type == IBAMR_FE_HEX27)
    {
        const int ne = type == IBAMR_FE_TET10 ? 6 : 12;
        double m = 0.0;
        for (int i = 0; i < ne; ++i)
        {
            const int* tr = type == IBAMR_FE_TET10 ? c_trip_tet10[i] : c_trip_hex27[i];
            const double l = sqrt(fe_d2(ndim, X, n[tr[0]], n[tr[1]])) + sqrt(fe_d2(ndim, X, n[tr[1]], n[tr[2]]));
            m = l > m ? l : m;
        }
        return m;
    }
    const int(*pr)[2];
    int np;
    switch (type)
    {
    case IBAMR_FE_TET4: pr = c_pairs_tet4; np = 6; break;
    case IBAMR_FE_HEX8:
    case IBAMR_FE_HEX20: pr = c_pairs_hex; np = 12; break;
    case IBAMR_FE_TRI3:
    case IBAMR_FE_TRI6: pr = c_pairs_tri; np = 3; break;
    case IBAMR_FE_QUAD4:
    case IBAMR_FE_QUAD8:
    case IBAMR_FE_QUAD9: pr = c_pairs_quad; np = 4; break;
    default: pr = c_pairs_tri; np = 1; break;  // EDGE2 / EDGE3: the two vertices
    }
    double m2 = 0.0;
    for (int i = 0; i < np; ++i)
    {
        const double l = fe_d2(ndim, X, n[pr[i][0]], n[pr[i][1]]);
        m2 = l > m2 ? l : m2;
    }
    return sqrt(m2);
}

// getQuadratureKey's adaptive order (libmesh_utilities.cpp:187-225): npts =
// int(ceil(point_density * hmax / dx_min)), at least 2 (first-order elements) or
// 3; QGAUSS order min(2 npts - 1, 43), QGRID npts.  An element with a node index
// outside [0, nnodes) gets order -1.
__global__ void k_fe_adaptive_orders(int type, int ndim, int nelem, int nbasis, int nnodes,
                                     const int* __restrict__ elem_nodes, const double* __restrict__ X, int quad_type,
                                     double point_density, double dx_min, int* __restrict__ order)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= nelem) return;
    const int* en = elem_nodes + (size_t)e * nbasis;
    for (int k = 0; k < nbasis; ++k)
        if (en[k] < 0 || en[k] >= nnodes)
        {
            order[e] = -1;
            return;
        }
    const bool first = type == IBAMR_FE_EDGE2 || type == IBAMR_FE_TRI3 || type == IBAMR_FE_QUAD4 ||
                       type == IBAMR_FE_TET4 || type == IBAMR_FE_HEX8;
    const double hmax = fe_max_edge(type, ndim, en, X);
    const double c = ceil(point_density * hmax / dx_min);
    int npts = c < 1.0e9 ? (int)c : 1000000000;  // (NaN too)
    if (npts < 3) npts = first ? 2 : 3;
    order[e] = quad_type == IBAMR_FE_QGAUSS ? min(2 * npts - 1, 43) : npts;
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

int ibamr_fe_quadrature_points_rules(int ndim, int nvars, int nelem, int nbasis, int nnodes, const int* elem_nodes,
                                     int nrules, const int* elem_rule, const int* rule_nqp, const long* rule_phi_off,
                                     const double* phi, const long* qp_off, long nqp_total, const double* JxW,
                                     const double* X_nodes, const double* F_nodes, double* X_qp, void* F_qp,
                                     int f_precision, void* stream)
{
    using ibamr::set_error;
    if (ndim != 2 && ndim != 3) return set_error(IBAMR_ERR_INVALID, "ndim must be 2 or 3");
    if (nelem < 0 || nbasis < 1 || nnodes < 0 || nrules < 1 || nqp_total < 0)
        return set_error(IBAMR_ERR_INVALID, "bad element or rule sizes");
    if (F_nodes && (nvars < 1 || nvars > FE_MAX_VARS))
        return set_error(IBAMR_ERR_INVALID, "nvars must be 1.." + std::to_string(FE_MAX_VARS));
    if (nqp_total > 0x7fffffffL) return set_error(IBAMR_ERR_CAPACITY, "too many quadrature points");
    if (nelem == 0 || nqp_total == 0) return IBAMR_OK;
    if (!elem_nodes || !elem_rule || !rule_nqp || !rule_phi_off || !phi || !qp_off || !X_nodes || !X_qp ||
        (F_nodes && !F_qp))
        return set_error(IBAMR_ERR_INVALID, "null argument");
    if (f_precision != IBAMR_FP64 && f_precision != IBAMR_FP32C) return set_error(IBAMR_ERR_INVALID, "bad precision");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256;
    const dim3 G((unsigned)((nqp_total + B - 1) / B));
    if (f_precision == IBAMR_FP64)
        hipLaunchKernelGGL(k_fe_qp_rules<double>, G, dim3(B), 0, s, ndim, nvars, nelem, nbasis, nnodes, nrules,
                           nqp_total, elem_nodes, elem_rule, rule_nqp, rule_phi_off, phi, qp_off, JxW, X_nodes,
                           F_nodes, X_qp, static_cast<double*>(F_qp));
    else
        hipLaunchKernelGGL(k_fe_qp_rules<float>, G, dim3(B), 0, s, ndim, nvars, nelem, nbasis, nnodes, nrules,
                           nqp_total, elem_nodes, elem_rule, rule_nqp, rule_phi_off, phi, qp_off, JxW, X_nodes,
                           F_nodes, X_qp, static_cast<float*>(F_qp));
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_qp_rules: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

int ibamr_fe_integrate_rhs_rules(const ibamr_fe_mesh* m, int nvars, int nrules, const int* elem_rule,
                                 const int* rule_nqp, const long* rule_phi_off, const double* phi, const long* qp_off,
                                 const double* JxW, const void* U_qp, int u_precision, double* F_nodes, int mode,
                                 void* stream)
{
    using ibamr::set_error;
    if (!m) return set_error(IBAMR_ERR_INVALID, "null mesh");
    if (nvars < 1 || nvars > FE_MAX_VARS) return set_error(IBAMR_ERR_INVALID, "nvars must be 1..8");
    if (nrules < 1) return set_error(IBAMR_ERR_INVALID, "nrules < 1");
    if (mode != IBAMR_SPREAD_ACCUMULATE && mode != IBAMR_SPREAD_OVERWRITE) return set_error(IBAMR_ERR_INVALID, "bad mode");
    if (u_precision != IBAMR_FP64 && u_precision != IBAMR_FP32C) return set_error(IBAMR_ERR_INVALID, "bad precision");
    if (m->nnodes == 0) return IBAMR_OK;
    if (!F_nodes || (m->nelem > 0 && (!elem_rule || !rule_nqp || !rule_phi_off || !phi || !qp_off || !JxW || !U_qp)))
        return set_error(IBAMR_ERR_INVALID, "null argument");
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int B = 256;
    const dim3 G((m->nnodes + B - 1) / B);
    const int ow = mode == IBAMR_SPREAD_OVERWRITE ? 1 : 0;
    if (u_precision == IBAMR_FP64)
        hipLaunchKernelGGL(k_fe_node_rhs_rules<double>, G, dim3(B), 0, s, nvars, m->nnodes, nrules, m->start, m->ent,
                           elem_rule, rule_nqp, rule_phi_off, phi, qp_off, JxW, static_cast<const double*>(U_qp),
                           F_nodes, ow);
    else
        hipLaunchKernelGGL(k_fe_node_rhs_rules<float>, G, dim3(B), 0, s, nvars, m->nnodes, nrules, m->start, m->ent,
                           elem_rule, rule_nqp, rule_phi_off, phi, qp_off, JxW, static_cast<const float*>(U_qp),
                           F_nodes, ow);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_node_rhs_rules: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

int ibamr_fe_nodal_mass(const ibamr_fe_mesh* m, int nqp, const double* phi, const double* JxW, double* M_diag,
                        void* stream)
{
    using ibamr::set_error;
    if (!m) return set_error(IBAMR_ERR_INVALID, "null mesh");
    if (nqp < 1) return set_error(IBAMR_ERR_INVALID, "nqp < 1");
    if (m->nnodes == 0) return IBAMR_OK;
    if (!M_diag || (m->nelem > 0 && (!phi || !JxW))) return set_error(IBAMR_ERR_INVALID, "null argument");
    const int B = 256;
    hipLaunchKernelGGL(k_fe_nodal_mass, dim3((m->nnodes + B - 1) / B), dim3(B), 0, static_cast<hipStream_t>(stream),
                       m->nnodes, nqp, m->start, m->ent, phi, JxW, M_diag);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_nodal_mass: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

int ibamr_fe_nodal_weight(int nnodes, int nvars, const void* V, int v_precision, const double* M_diag, void* out,
                          int out_precision, int mode, void* stream)
{
    using ibamr::set_error;
    if (nnodes < 0 || nvars < 1) return set_error(IBAMR_ERR_INVALID, "bad sizes");
    if (mode != IBAMR_SPREAD_ACCUMULATE && mode != IBAMR_SPREAD_OVERWRITE) return set_error(IBAMR_ERR_INVALID, "bad mode");
    if ((v_precision != IBAMR_FP64 && v_precision != IBAMR_FP32C) ||
        (out_precision != IBAMR_FP64 && out_precision != IBAMR_FP32C))
        return set_error(IBAMR_ERR_INVALID, "bad precision");
    if (nnodes == 0) return IBAMR_OK;
    if (!V || !M_diag || !out) return set_error(IBAMR_ERR_INVALID, "null argument");
    const long n = (long)nnodes * nvars;
    const int B = 256;
    const dim3 G((unsigned)((n + B - 1) / B));
    hipStream_t s = static_cast<hipStream_t>(stream);
    const int ow = mode == IBAMR_SPREAD_OVERWRITE ? 1 : 0;
#define NW(TV, TO)                                                                                               \
    hipLaunchKernelGGL((k_fe_nodal_weight<TV, TO>), G, dim3(B), 0, s, (long)nnodes, nvars,                      \
                       static_cast<const TV*>(V), M_diag, static_cast<TO*>(out), ow)
    if (v_precision == IBAMR_FP64)
    {
        if (out_precision == IBAMR_FP64) NW(double, double);
        else NW(double, float);
    }
    else
    {
        if (out_precision == IBAMR_FP64) NW(float, double);
        else NW(float, float);
    }
#undef NW
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_nodal_weight: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

int ibamr_fe_adaptive_quadrature_orders(int elem_type, int ndim, int nelem, int nbasis, int nnodes,
                                        const int* elem_nodes, const double* X_nodes, int quad_type,
                                        double point_density, double dx_min, int* order, void* stream)
{
    using ibamr::set_error;
    static const int nb_min[12] = {2, 3, 3, 6, 4, 8, 9, 4, 10, 8, 20, 27};
    if (elem_type < IBAMR_FE_EDGE2 || elem_type > IBAMR_FE_HEX27)
        return set_error(IBAMR_ERR_INVALID, "unknown element type " + std::to_string(elem_type));
    if (ndim < 1 || ndim > 3) return set_error(IBAMR_ERR_INVALID, "ndim must be 1..3");
    if (nbasis < nb_min[elem_type])
        return set_error(IBAMR_ERR_INVALID, "nbasis " + std::to_string(nbasis) + " is below the element type's " +
                                                std::to_string(nb_min[elem_type]) + " nodes");
    if (quad_type != IBAMR_FE_QGAUSS && quad_type != IBAMR_FE_QGRID)
        return set_error(IBAMR_ERR_INVALID, "adaptive quadrature rules are available only for QGAUSS or QGRID");
    if (!(dx_min > 0.0)) return set_error(IBAMR_ERR_INVALID, "dx_min must be positive");
    if (nelem < 0 || nnodes < 0) return set_error(IBAMR_ERR_INVALID, "bad sizes");
    if (nelem == 0) return IBAMR_OK;
    if (!elem_nodes || !X_nodes || !order) return set_error(IBAMR_ERR_INVALID, "null argument");
    const int B = 256;
    hipLaunchKernelGGL(k_fe_adaptive_orders, dim3((nelem + B - 1) / B), dim3(B), 0, static_cast<hipStream_t>(stream),
                       elem_type, ndim, nelem, nbasis, nnodes, elem_nodes, X_nodes, quad_type, point_density, dx_min,
                       order);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess)
        return set_error(IBAMR_ERR_DEVICE, std::string("k_fe_adaptive_orders: ") + hipGetErrorString(e));
    return IBAMR_OK;
}

} // extern "C"
