// Host-side model construction of the sparlab_hip library: structured grid,
// DOF map, the cached element stiffness, the K_ff pattern and the density
// filter. All of it runs once per problem; the per-iteration work is on the
// device.
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <map>
#include <sstream>
#include <tuple>

#include "slx_model.hpp"

namespace slx {

namespace {
const double kG = 1.0 / std::sqrt(3.0);
const double kGauss2[2] = {-kG, kG};
// Hex8 natural corner coordinates in VTK order (src/elements/Hex8.cpp:15-17)
const double kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

// ke += scale * B^T D B with B nv x nd (row-major) and D nv x nv; the
// products are formed as (B^T D) then (B^T D) B, each inner sum in index order.
void add_btdb(const double* b, const double* d, int nv, int nd, double scale, double* ke) {
  std::vector<double> bd((size_t)nd * nv);
  for (int p = 0; p < nd; ++p)
    for (int q = 0; q < nv; ++q) {
      double s = 0.0;
      for (int r = 0; r < nv; ++r) s += b[r * nd + p] * d[r * nv + q];
      bd[(size_t)p * nv + q] = s;
    }
  for (int p = 0; p < nd; ++p)
    for (int t = 0; t < nd; ++t) {
      double s = 0.0;
      for (int q = 0; q < nv; ++q) s += bd[(size_t)p * nv + q] * b[q * nd + t];
      ke[p * nd + t] += scale * s;
    }
}

void symmetrise(double* ke, int nd) {
  std::vector<double> out((size_t)nd * nd);
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) out[(size_t)i * nd + j] = 0.5 * (ke[i * nd + j] + ke[j * nd + i]);
  std::copy(out.begin(), out.end(), ke);
}
}  // namespace

// Hex8 stiffness of a box cell of sides (hx, hy, hz): 2x2x2 Gauss, explicit
// cofactor Jacobian inverse, Ke = sum detJ w B^T D B, symmetrised
// (src/elements/Hex8.cpp:46-101,172-192; src/material/IsotropicMaterial.cpp:56-66).
std::vector<double> hex8_ke0(const double* xe /* 3 x 8 */, double E, double nu) {
  double d[36] = {0};
  const double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double g = E / (2.0 * (1.0 + nu));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) d[i * 6 + j] = lambda;
    d[i * 6 + i] = lambda + 2.0 * g;
    d[(3 + i) * 6 + 3 + i] = g;
  }
  std::vector<double> ke(576, 0.0);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        const double xi = kGauss2[i], eta = kGauss2[j], zeta = kGauss2[k];
        const double w = 1.0 * 1.0 * 1.0;
        double gr[8][3];
        for (int a = 0; a < 8; ++a) {
          const double fx = 1.0 + xi * kCorner[a][0], fy = 1.0 + eta * kCorner[a][1], fz = 1.0 + zeta * kCorner[a][2];
          gr[a][0] = 0.125 * kCorner[a][0] * fy * fz;
          gr[a][1] = 0.125 * kCorner[a][1] * fx * fz;
          gr[a][2] = 0.125 * kCorner[a][2] * fx * fy;
        }
        double J[3][3];
        for (int r = 0; r < 3; ++r)
          for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += xe[r * 8 + a] * gr[a][c];
            J[r][c] = s;
          }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        const double det = J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02;
        if (!(det > 0.0)) fail(SLX_MODEL_ERROR, "Hex8 Jacobian determinant is not positive");
        double ji[3][3];
        ji[0][0] = c00 / det;
        ji[1][0] = c01 / det;
        ji[2][0] = c02 / det;
        ji[0][1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) / det;
        ji[1][1] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) / det;
        ji[2][1] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) / det;
        ji[0][2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) / det;
        ji[1][2] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) / det;
        ji[2][2] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) / det;
        double b[6 * 24] = {0};
        for (int a = 0; a < 8; ++a) {
          double dn[3];
          for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += gr[a][q] * ji[q][c];
            dn[c] = s;
          }
          const int col = 3 * a;
          b[0 * 24 + col + 0] = dn[0];
          b[1 * 24 + col + 1] = dn[1];
          b[2 * 24 + col + 2] = dn[2];
          b[3 * 24 + col + 0] = dn[1];
          b[3 * 24 + col + 1] = dn[0];
          b[4 * 24 + col + 1] = dn[2];
          b[4 * 24 + col + 2] = dn[1];
          b[5 * 24 + col + 0] = dn[2];
          b[5 * 24 + col + 2] = dn[0];
        }
        add_btdb(b, d, 6, 24, det * w, ke.data());
      }
  symmetrise(ke.data(), 24);
  return ke;
}

// Q4 plane-stress stiffness (src/elements/Quad4.cpp:44-81,117-137;
// src/material/IsotropicMaterial.cpp:34-43).
std::vector<double> quad4_ke0(const double* xe /* 2 x 4 */, double E, double nu, double t) {
  if (!(t > 0.0)) fail(SLX_CONFIG_ERROR, "element thickness must be positive");
  double d[9] = {0};
  const double f = E / (1.0 - nu * nu);
  d[0] = f;
  d[1] = f * nu;
  d[3] = f * nu;
  d[4] = f;
  d[8] = f * (1.0 - nu) / 2.0;
  std::vector<double> ke(64, 0.0);
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) {
      const double xi = kGauss2[i], eta = kGauss2[j];
      const double w = 1.0 * 1.0;
      const double g[4][2] = {{-0.25 * (1.0 - eta), -0.25 * (1.0 - xi)},
                              {0.25 * (1.0 - eta), -0.25 * (1.0 + xi)},
                              {0.25 * (1.0 + eta), 0.25 * (1.0 + xi)},
                              {-0.25 * (1.0 + eta), 0.25 * (1.0 - xi)}};
      double J[2][2];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
          double s = 0.0;
          for (int a = 0; a < 4; ++a) s += xe[r * 4 + a] * g[a][c];
          J[r][c] = s;
        }
      const double det = J[0][0] * J[1][1] - J[0][1] * J[1][0];
      if (!(det > 0.0)) fail(SLX_MODEL_ERROR, "Quad4 Jacobian determinant is not positive");
      const double ji00 = J[1][1] / det, ji01 = -J[0][1] / det, ji10 = -J[1][0] / det, ji11 = J[0][0] / det;
      double b[24] = {0};
      for (int a = 0; a < 4; ++a) {
        const double dx = g[a][0] * ji00 + g[a][1] * ji10;
        const double dy = g[a][0] * ji01 + g[a][1] * ji11;
        b[0 * 8 + 2 * a] = dx;
        b[1 * 8 + 2 * a + 1] = dy;
        b[2 * 8 + 2 * a] = dy;
        b[2 * 8 + 2 * a + 1] = dx;
      }
      add_btdb(b, d, 3, 8, t * det * w, ke.data());
    }
  symmetrise(ke.data(), 8);
  return ke;
}

void build_structured_grid(Context& c) {
  const int nnx = c.nx + 1, nny = c.ny + 1, nnz = c.dim == 3 ? c.nz + 1 : 1;
  const double dx = c.lx / double(c.nx), dy = c.ly / double(c.ny), dz = c.dim == 3 ? c.lz / double(c.nz) : 0.0;
  c.nnodes = (long long)nnx * nny * nnz;
  c.nelem = (long long)c.nx * c.ny * (c.dim == 3 ? c.nz : 1);
  c.ndpn = c.dim;
  c.npe = c.dim == 3 ? 8 : 4;
  c.ndofs = c.nnodes * c.dim;
  c.coords.resize((size_t)c.nnodes * c.dim);
  for (int k = 0; k < nnz; ++k)
    for (int j = 0; j < nny; ++j)
      for (int i = 0; i < nnx; ++i) {
        const size_t n = ((size_t)k * nny + j) * nnx + i;
        c.coords[n * c.dim + 0] = 0.0 + double(i) * dx;
        c.coords[n * c.dim + 1] = 0.0 + double(j) * dy;
        if (c.dim == 3) c.coords[n * c.dim + 2] = 0.0 + double(k) * dz;
      }
}

// element 0's corner coordinates, dim x npe
std::vector<double> element0_coords(const Context& c) {
  const int nnx = c.nx + 1, nny = c.ny + 1;
  std::vector<int> nodes;
  if (c.dim == 3) {
    const int lay = nnx * nny;
    nodes = {0, 1, nnx + 1, nnx, lay, lay + 1, lay + nnx + 1, lay + nnx};
  } else {
    nodes = {0, 1, nnx + 1, nnx};
  }
  std::vector<double> xe((size_t)c.dim * c.npe);
  for (int a = 0; a < c.npe; ++a)
    for (int i = 0; i < c.dim; ++i) xe[(size_t)i * c.npe + a] = c.coords[(size_t)nodes[(size_t)a] * c.dim + i];
  return xe;
}

double element_volume(const Context& c) {
  // every cell of a uniform grid has the measure of element 0: the shoelace
  // area times the thickness in 2-D, the 2x2x2 Gauss volume in 3-D
  // (src/mesh/Mesh.cpp:377-396, src/fem/FemModel.cpp:236-240,422-427)
  const std::vector<double> xe = element0_coords(c);
  if (c.dim == 2) {
    double twice = 0.0;
    for (int a = 0; a < 4; ++a) {
      const int b = (a + 1) % 4;
      twice += xe[0 * 4 + a] * xe[1 * 4 + b] - xe[0 * 4 + b] * xe[1 * 4 + a];
    }
    return 0.5 * twice * c.thickness;
  }
  double vol = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        double gr[8][3];
        for (int a = 0; a < 8; ++a) {
          const double fx = 1.0 + kGauss2[i] * kCorner[a][0], fy = 1.0 + kGauss2[j] * kCorner[a][1],
                       fz = 1.0 + kGauss2[k] * kCorner[a][2];
          gr[a][0] = 0.125 * kCorner[a][0] * fy * fz;
          gr[a][1] = 0.125 * kCorner[a][1] * fx * fz;
          gr[a][2] = 0.125 * kCorner[a][2] * fx * fy;
        }
        double J[3][3];
        for (int r = 0; r < 3; ++r)
          for (int cc = 0; cc < 3; ++cc) {
            double s = 0.0;
            for (int a = 0; a < 8; ++a) s += xe[(size_t)r * 8 + a] * gr[a][cc];
            J[r][cc] = s;
          }
        const double c00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
        const double c01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
        const double c02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
        vol += 1.0 * (J[0][0] * c00 + J[0][1] * c01 + J[0][2] * c02);
      }
  return vol;
}

// K_ff pattern: the row of free DOF (I, ci) lists the free DOFs of the nodes
// sharing a cell with I, ascending (the node neighbourhoods of
// src/fem/Assembler.cpp:111-155 reduced by src/fem/Assembler.cpp:415-436).
HCsr structured_kff_pattern(const Context& c) {
  const int nf = (int)c.free_dofs.size();
  const int nnx = c.nx + 1, nny = c.ny + 1, nnz = c.dim == 3 ? c.nz + 1 : 1;
  const int dz0 = c.dim == 3 ? -1 : 0, dz1 = c.dim == 3 ? 1 : 0;
  HCsr p;
  p.rows = p.cols = nf;
  p.ptr.assign((size_t)nf + 1, 0);
  auto visit = [&](int row, auto&& emit) {
    const int g = c.free_dofs[(size_t)row];
    const int I = g / c.dim;
    const int ix = I % nnx, iy = (I / nnx) % nny, iz = I / (nnx * nny);
    for (int dz = dz0; dz <= dz1; ++dz) {
      const int jz = iz + dz;
      if (jz < 0 || jz >= nnz) continue;
      for (int dy = -1; dy <= 1; ++dy) {
        const int jy = iy + dy;
        if (jy < 0 || jy >= nny) continue;
        for (int dx = -1; dx <= 1; ++dx) {
          const int jx = ix + dx;
          if (jx < 0 || jx >= nnx) continue;
          const int J = (jz * nny + jy) * nnx + jx;
          for (int cj = 0; cj < c.dim; ++cj) {
            const int r = c.reduced[(size_t)J * c.dim + cj];
            if (r >= 0) emit(r);
          }
        }
      }
    }
  };
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nf; ++i) {
    int cnt = 0;
    visit(i, [&](int) { ++cnt; });
    p.ptr[(size_t)i + 1] = cnt;
  }
  for (int i = 0; i < nf; ++i) {
    const long long nx2 = (long long)p.ptr[(size_t)i] + p.ptr[(size_t)i + 1];
    if (nx2 > 2147483647LL) fail(SLX_MODEL_ERROR, "K_ff exceeds 2^31 stored entries");
    p.ptr[(size_t)i + 1] = (int)nx2;
  }
  p.col.resize((size_t)p.ptr.back());
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nf; ++i) {
    int q = p.ptr[(size_t)i];
    visit(i, [&](int r) { p.col[(size_t)q++] = r; });
  }
  return p;
}

// Density filter (src/topopt/DensityFilter.cpp:19-74,115-173): weights
// r - |c_e - c_i| > 0 over the 3^dim buckets (cell size = radius) around each
// centroid, visited z slowest, then y, then x, bucket contents ascending;
// row sums in that order; H_ei = w_ei / rowsum_e stored with ascending columns.
void build_filter(Context& c, double radius) {
  if (!(radius > 0.0)) fail(SLX_CONFIG_ERROR, "filter radius must be positive");
  const int dim = c.dim;
  const long long ne = c.nelem;
  const int nnx = c.nx + 1, nny = c.ny + 1;
  std::vector<double> cen((size_t)ne * 3, 0.0);
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < ne; ++e) {
    int nodes[8];
    if (dim == 3) {
      const int ex = (int)(e % c.nx), ey = (int)((e / c.nx) % c.ny), ez = (int)(e / ((long long)c.nx * c.ny));
      const int n0 = (ez * nny + ey) * nnx + ex, lay = nnx * nny;
      const int t[8] = {n0, n0 + 1, n0 + nnx + 1, n0 + nnx, n0 + lay, n0 + lay + 1, n0 + lay + nnx + 1, n0 + lay + nnx};
      std::copy(t, t + 8, nodes);
    } else {
      const int ex = (int)(e % c.nx), ey = (int)(e / c.nx);
      const int n0 = ey * nnx + ex;
      const int t[4] = {n0, n0 + 1, n0 + nnx + 1, n0 + nnx};
      std::copy(t, t + 4, nodes);
    }
    double s[3] = {0, 0, 0};
    for (int a = 0; a < c.npe; ++a)
      for (int d = 0; d < dim; ++d) s[d] += c.coords[(size_t)nodes[a] * dim + d];
    for (int d = 0; d < dim; ++d) cen[(size_t)e * 3 + d] = s[d] / double(c.npe);
  }
  auto cell = [&](double v) { return (long)std::floor(v / radius); };
  std::map<std::tuple<long, long, long>, std::vector<int>> buckets;
  for (long long e = 0; e < ne; ++e)
    buckets[{cell(cen[(size_t)e * 3]), cell(cen[(size_t)e * 3 + 1]), dim == 3 ? cell(cen[(size_t)e * 3 + 2]) : 0}]
        .push_back((int)e);
  std::vector<std::vector<std::pair<int, double>>> rows((size_t)ne);
  bool empty_row = false;
#pragma omp parallel for schedule(dynamic, 1024)
  for (long long e = 0; e < ne; ++e) {
    const double* ce = &cen[(size_t)e * 3];
    const long ix = cell(ce[0]), iy = cell(ce[1]), iz = dim == 3 ? cell(ce[2]) : 0;
    const long dzr = dim == 3 ? 1 : 0;
    double sum = 0.0;
    auto& out = rows[(size_t)e];
    for (long dz = -dzr; dz <= dzr; ++dz)
      for (long dy = -1; dy <= 1; ++dy)
        for (long dx = -1; dx <= 1; ++dx) {
          auto it = buckets.find({ix + dx, iy + dy, iz + dz});
          if (it == buckets.end()) continue;
          for (int i : it->second) {
            double s2 = 0.0;
            for (int d = 0; d < dim; ++d) {
              const double t = ce[d] - cen[(size_t)i * 3 + d];
              s2 += t * t;
            }
            const double w = radius - std::sqrt(s2);
            if (w <= 0.0) continue;
            out.emplace_back(i, w);
            sum += w;
          }
        }
    if (!(sum > 0.0)) {
      empty_row = true;
      continue;
    }
    std::sort(out.begin(), out.end());
    for (auto& pr : out) pr.second = pr.second / sum;
  }
  if (empty_row) fail(SLX_CONFIG_ERROR, "a filter row is empty; the mesh may contain non-finite centroids");
  HCsr& H = c.H;
  H.rows = H.cols = (int)ne;
  H.ptr.assign((size_t)ne + 1, 0);
  for (long long e = 0; e < ne; ++e) H.ptr[(size_t)e + 1] = H.ptr[(size_t)e] + (int)rows[(size_t)e].size();
  H.col.resize((size_t)H.ptr.back());
  H.val.resize((size_t)H.ptr.back());
#pragma omp parallel for schedule(static)
  for (long long e = 0; e < ne; ++e) {
    int q = H.ptr[(size_t)e];
    for (auto& pr : rows[(size_t)e]) {
      H.col[(size_t)q] = pr.first;
      H.val[(size_t)q] = pr.second;
      ++q;
    }
  }
  // transpose: row j lists H_ej for ascending e (the order of H^T v in
  // src/topopt/DensityFilter.cpp:186-195)
  HCsr& T = c.Ht;
  T.rows = T.cols = (int)ne;
  T.ptr.assign((size_t)ne + 1, 0);
  for (long long k = 0; k < H.nnz(); ++k) ++T.ptr[(size_t)H.col[(size_t)k] + 1];
  for (long long j = 0; j < ne; ++j) T.ptr[(size_t)j + 1] += T.ptr[(size_t)j];
  T.col.resize((size_t)H.nnz());
  T.val.resize((size_t)H.nnz());
  std::vector<int> next(T.ptr.begin(), T.ptr.end() - 1);
  for (long long e = 0; e < ne; ++e)
    for (int k = H.ptr[(size_t)e]; k < H.ptr[(size_t)e + 1]; ++k) {
      const int q = next[(size_t)H.col[(size_t)k]]++;
      T.col[(size_t)q] = (int)e;
      T.val[(size_t)q] = H.val[(size_t)k];
    }
  c.filter_on = true;
  c.filter_radius = radius;
}

void identity_filter(Context& c) {
  const int ne = (int)c.nelem;
  for (HCsr* h : {&c.H, &c.Ht}) {
    h->rows = h->cols = ne;
    h->ptr.resize((size_t)ne + 1);
    h->col.resize((size_t)ne);
    h->val.assign((size_t)ne, 1.0);
    for (int e = 0; e <= ne; ++e) h->ptr[(size_t)e] = e;
    for (int e = 0; e < ne; ++e) h->col[(size_t)e] = e;
  }
  c.filter_on = false;
  c.filter_radius = 0.0;
}

}  // namespace slx
