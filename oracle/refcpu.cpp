// refcpu: CPU restatement of SparLab's SIMP linear-solve path (test oracle and
// CPU baseline only; never linked by the product). Citations are into the
// SparLab source tree.
#include "refcpu.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <sstream>
#include <tuple>

#if defined(_OPENMP)
#include <omp.h>
#endif

namespace refcpu {
namespace {

// Multigrid.cpp:22-25: fixed 4096-element chunks make inner products
// independent of the thread count; loops below 16384 rows stay serial.
constexpr Index kChunk = 4096;
constexpr Index kParallelRows = 16384;

[[noreturn]] void fail(int code, const std::string& m) { throw Error(code, m); }

// Gauss-Legendre 2-point rule (Quadrature.cpp:19-20).
const double kGauss2[2] = {-1.0 / std::sqrt(3.0), 1.0 / std::sqrt(3.0)};

// Natural coordinates of the Hex8 corners in VTK order (Hex8.cpp:15-17).
const double kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

}  // namespace

// ---------------------------------------------------------------------------
// Mesh
// ---------------------------------------------------------------------------
Mesh structured_quad(Index nx, Index ny, double lx, double ly) {
  // StructuredMesh.cpp:157-194
  if (nx < 1 || ny < 1 || !(lx > 0) || !(ly > 0)) fail(1, "structured quad mesh needs positive counts and lengths");
  Mesh m;
  m.dim = 2;
  m.npe = 4;
  m.nx = nx;
  m.ny = ny;
  m.lx = lx;
  m.ly = ly;
  const Index nnx = nx + 1, nny = ny + 1;
  const double dx = lx / double(nx), dy = ly / double(ny);
  m.nn = nnx * nny;
  m.coords.resize((size_t)m.nn * 2);
  for (Index j = 0; j < nny; ++j)
    for (Index i = 0; i < nnx; ++i) {
      const Index n = j * nnx + i;
      m.coords[(size_t)n * 2 + 0] = 0.0 + double(i) * dx;
      m.coords[(size_t)n * 2 + 1] = 0.0 + double(j) * dy;
    }
  m.ne = nx * ny;
  m.conn.reserve((size_t)m.ne * 4);
  for (Index j = 0; j < ny; ++j)
    for (Index i = 0; i < nx; ++i) {
      const Index n0 = j * nnx + i;
      const Index c[4] = {n0, n0 + 1, n0 + nnx + 1, n0 + nnx};
      m.conn.insert(m.conn.end(), c, c + 4);
    }
  return m;
}

Mesh structured_hex(Index nx, Index ny, Index nz, double lx, double ly, double lz) {
  // StructuredMesh.cpp:196-247
  if (nx < 1 || ny < 1 || nz < 1 || !(lx > 0) || !(ly > 0) || !(lz > 0))
    fail(1, "structured hex mesh needs positive counts and lengths");
  Mesh m;
  m.dim = 3;
  m.npe = 8;
  m.nx = nx;
  m.ny = ny;
  m.nz = nz;
  m.lx = lx;
  m.ly = ly;
  m.lz = lz;
  const Index nnx = nx + 1, nny = ny + 1, nnz = nz + 1;
  const double dx = lx / double(nx), dy = ly / double(ny), dz = lz / double(nz);
  m.nn = nnx * nny * nnz;
  m.coords.resize((size_t)m.nn * 3);
  for (Index k = 0; k < nnz; ++k)
    for (Index j = 0; j < nny; ++j)
      for (Index i = 0; i < nnx; ++i) {
        const Index n = k * nnx * nny + j * nnx + i;
        m.coords[(size_t)n * 3 + 0] = 0.0 + double(i) * dx;
        m.coords[(size_t)n * 3 + 1] = 0.0 + double(j) * dy;
        m.coords[(size_t)n * 3 + 2] = 0.0 + double(k) * dz;
      }
  m.ne = nx * ny * nz;
  m.conn.reserve((size_t)m.ne * 8);
  const Index layer = nnx * nny;
  for (Index k = 0; k < nz; ++k)
    for (Index j = 0; j < ny; ++j)
      for (Index i = 0; i < nx; ++i) {
        const Index n0 = k * layer + j * nnx + i;
        const Index c[8] = {n0,         n0 + 1,         n0 + nnx + 1,         n0 + nnx,
                            n0 + layer, n0 + layer + 1, n0 + layer + nnx + 1, n0 + layer + nnx};
        m.conn.insert(m.conn.end(), c, c + 8);
      }
  return m;
}

void element_centroid(const Mesh& m, Index e, double* c) {
  double s[3] = {0, 0, 0};
  const Index* nodes = m.element_nodes(e);
  for (int a = 0; a < m.npe; ++a)
    for (int d = 0; d < m.dim; ++d) s[d] += m.coords[(size_t)nodes[a] * m.dim + d];
  for (int d = 0; d < 3; ++d) c[d] = d < m.dim ? s[d] / double(m.npe) : 0.0;
}

namespace {
void hex8_gradients(double xi, double eta, double zeta, double g[8][3]) {
  // Hex8.cpp:114-126
  for (int a = 0; a < 8; ++a) {
    const double fx = 1.0 + xi * kCorner[a][0];
    const double fy = 1.0 + eta * kCorner[a][1];
    const double fz = 1.0 + zeta * kCorner[a][2];
    g[a][0] = 0.125 * kCorner[a][0] * fy * fz;
    g[a][1] = 0.125 * kCorner[a][1] * fx * fz;
    g[a][2] = 0.125 * kCorner[a][2] * fx * fy;
  }
}
void quad4_gradients(double xi, double eta, double g[4][2]) {
  // Quad4.cpp:94-102
  g[0][0] = -0.25 * (1.0 - eta);
  g[0][1] = -0.25 * (1.0 - xi);
  g[1][0] = 0.25 * (1.0 - eta);
  g[1][1] = -0.25 * (1.0 + xi);
  g[2][0] = 0.25 * (1.0 + eta);
  g[2][1] = 0.25 * (1.0 + xi);
  g[3][0] = -0.25 * (1.0 + eta);
  g[3][1] = 0.25 * (1.0 - xi);
}
// J(i,j) = sum_a x(i,a) g(a,j)
void hex8_jacobian(const double* xe, const double g[8][3], double jac[3][3]) {
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
      for (int a = 0; a < 8; ++a) s += xe[i * 8 + a] * g[a][j];
      jac[i][j] = s;
    }
}
double det3(const double j[3][3]) {
  const double c00 = j[1][1] * j[2][2] - j[1][2] * j[2][1];
  const double c01 = j[1][2] * j[2][0] - j[1][0] * j[2][2];
  const double c02 = j[1][0] * j[2][1] - j[1][1] * j[2][0];
  return j[0][0] * c00 + j[0][1] * c01 + j[0][2] * c02;
}
}  // namespace

double element_measure(const Mesh& m, Index e) {
  const Index* nodes = m.element_nodes(e);
  if (m.dim == 2) {  // Mesh.cpp:377-386, shoelace
    double twice = 0.0;
    for (int a = 0; a < 4; ++a) {
      const double* p = &m.coords[(size_t)nodes[a] * 2];
      const double* q = &m.coords[(size_t)nodes[(a + 1) % 4] * 2];
      twice += p[0] * q[1] - q[0] * p[1];
    }
    return 0.5 * twice;
  }
  double xe[24];  // Hex8.cpp:128-142
  for (int a = 0; a < 8; ++a)
    for (int i = 0; i < 3; ++i) xe[i * 8 + a] = m.coords[(size_t)nodes[a] * 3 + i];
  double volume = 0.0;
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        double g[8][3], jac[3][3];
        hex8_gradients(kGauss2[i], kGauss2[j], kGauss2[k], g);
        hex8_jacobian(xe, g, jac);
        volume += 1.0 * det3(jac);
      }
  return volume;
}

double mean_element_size(const Mesh& m) {  // Mesh.cpp:398-406
  double sum = 0.0;
  for (Index e = 0; e < m.ne; ++e) sum += element_measure(m, e);
  const double mean = sum / double(m.ne);
  return m.dim == 2 ? std::sqrt(mean) : std::cbrt(mean);
}

// ---------------------------------------------------------------------------
// Material, element stiffness
// ---------------------------------------------------------------------------
void d_plane_stress(double E, double nu, double d[9]) {
  std::fill(d, d + 9, 0.0);
  const double f = E / (1.0 - nu * nu);
  d[0] = f;
  d[1] = f * nu;
  d[3] = f * nu;
  d[4] = f;
  d[8] = f * (1.0 - nu) / 2.0;
}

void d_three_dimensional(double E, double nu, double d[36]) {
  std::fill(d, d + 36, 0.0);
  const double lambda = E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
  const double g = E / (2.0 * (1.0 + nu));
  for (int i = 0; i < 3; ++i) {
    for (int j = 0; j < 3; ++j) d[i * 6 + j] = lambda;
    d[i * 6 + i] = lambda + 2.0 * g;
    d[(3 + i) * 6 + 3 + i] = g;
  }
}

namespace {
// ke(p,s) += scale * sum_q (sum_r b(r,p) d(r,q)) b(q,s)
void add_btdb(const double* b, const double* d, int nv, int nd, double scale, double* ke) {
  std::vector<double> tmp((size_t)nd * nv);
  for (int p = 0; p < nd; ++p)
    for (int q = 0; q < nv; ++q) {
      double s = 0.0;
      for (int r = 0; r < nv; ++r) s += b[r * nd + p] * d[r * nv + q];
      tmp[(size_t)p * nv + q] = s;
    }
  for (int p = 0; p < nd; ++p)
    for (int s2 = 0; s2 < nd; ++s2) {
      double s = 0.0;
      for (int q = 0; q < nv; ++q) s += tmp[(size_t)p * nv + q] * b[q * nd + s2];
      ke[p * nd + s2] += scale * s;
    }
}
void symmetrise(double* ke, int nd) {
  std::vector<double> out((size_t)nd * nd);
  for (int i = 0; i < nd; ++i)
    for (int j = 0; j < nd; ++j) out[(size_t)i * nd + j] = 0.5 * (ke[i * nd + j] + ke[j * nd + i]);
  std::copy(out.begin(), out.end(), ke);
}
}  // namespace

void hex8_stiffness(const double* xe, const double* d, double* ke) {
  std::fill(ke, ke + 576, 0.0);
  for (int k = 0; k < 2; ++k)
    for (int j = 0; j < 2; ++j)
      for (int i = 0; i < 2; ++i) {
        const double w = 1.0 * 1.0 * 1.0;
        double g[8][3], jac[3][3];
        hex8_gradients(kGauss2[i], kGauss2[j], kGauss2[k], g);
        hex8_jacobian(xe, g, jac);
        // Hex8.cpp:53-74: explicit cofactor inverse
        const double c00 = jac[1][1] * jac[2][2] - jac[1][2] * jac[2][1];
        const double c01 = jac[1][2] * jac[2][0] - jac[1][0] * jac[2][2];
        const double c02 = jac[1][0] * jac[2][1] - jac[1][1] * jac[2][0];
        const double det = jac[0][0] * c00 + jac[0][1] * c01 + jac[0][2] * c02;
        if (!(det > 0.0)) fail(4, "Hex8 Jacobian determinant is not positive");
        double ji[3][3];
        ji[0][0] = c00 / det;
        ji[1][0] = c01 / det;
        ji[2][0] = c02 / det;
        ji[0][1] = (jac[0][2] * jac[2][1] - jac[0][1] * jac[2][2]) / det;
        ji[1][1] = (jac[0][0] * jac[2][2] - jac[0][2] * jac[2][0]) / det;
        ji[2][1] = (jac[0][1] * jac[2][0] - jac[0][0] * jac[2][1]) / det;
        ji[0][2] = (jac[0][1] * jac[1][2] - jac[0][2] * jac[1][1]) / det;
        ji[1][2] = (jac[0][2] * jac[1][0] - jac[0][0] * jac[1][2]) / det;
        ji[2][2] = (jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0]) / det;
        double b[6 * 24] = {0};
        for (int a = 0; a < 8; ++a) {  // Hex8.cpp:83-101
          double dn[3];
          for (int c = 0; c < 3; ++c) {
            double s = 0.0;
            for (int q = 0; q < 3; ++q) s += g[a][q] * ji[q][c];
            dn[c] = s;
          }
          const int c = 3 * a;
          b[0 * 24 + c + 0] = dn[0];
          b[1 * 24 + c + 1] = dn[1];
          b[2 * 24 + c + 2] = dn[2];
          b[3 * 24 + c + 0] = dn[1];
          b[3 * 24 + c + 1] = dn[0];
          b[4 * 24 + c + 1] = dn[2];
          b[4 * 24 + c + 2] = dn[1];
          b[5 * 24 + c + 0] = dn[2];
          b[5 * 24 + c + 2] = dn[0];
        }
        add_btdb(b, d, 6, 24, det * w, ke);
      }
  symmetrise(ke, 24);
}

void quad4_stiffness(const double* xe, const double* d, double t, double* ke) {
  if (!(t > 0.0)) fail(1, "element thickness must be positive");
  std::fill(ke, ke + 64, 0.0);
  for (int j = 0; j < 2; ++j)
    for (int i = 0; i < 2; ++i) {
      const double w = 1.0 * 1.0;
      double g[4][2];
      quad4_gradients(kGauss2[i], kGauss2[j], g);
      double jac[2][2];
      for (int r = 0; r < 2; ++r)
        for (int c = 0; c < 2; ++c) {
          double s = 0.0;
          for (int a = 0; a < 4; ++a) s += xe[r * 4 + a] * g[a][c];
          jac[r][c] = s;
        }
      const double det = jac[0][0] * jac[1][1] - jac[0][1] * jac[1][0];
      if (!(det > 0.0)) fail(4, "Quad4 Jacobian determinant is not positive");
      double ji[2][2];
      ji[0][0] = jac[1][1] / det;
      ji[0][1] = -jac[0][1] / det;
      ji[1][0] = -jac[1][0] / det;
      ji[1][1] = jac[0][0] / det;
      double b[3 * 8] = {0};
      for (int a = 0; a < 4; ++a) {  // Quad4.cpp:72-79
        const double dx = g[a][0] * ji[0][0] + g[a][1] * ji[1][0];
        const double dy = g[a][0] * ji[0][1] + g[a][1] * ji[1][1];
        b[0 * 8 + 2 * a] = dx;
        b[1 * 8 + 2 * a + 1] = dy;
        b[2 * 8 + 2 * a] = dy;
        b[2 * 8 + 2 * a + 1] = dx;
      }
      add_btdb(b, d, 3, 8, t * det * w, ke);
    }
  symmetrise(ke, 8);
}

// ---------------------------------------------------------------------------
// DOFs, pattern, scatter, reduction
// ---------------------------------------------------------------------------
void Dofs::rebuild() {  // DofManager.cpp:58-71
  free.clear();
  reduced.assign((size_t)ndofs, -1);
  for (Index d = 0; d < ndofs; ++d) {
    if (fixed[(size_t)d]) continue;
    reduced[(size_t)d] = (Index)free.size();
    free.push_back(d);
  }
}

Pattern build_pattern(const Mesh& m, int dim) {  // Assembler.cpp:85-172
  const Index nn = m.nn, ne = m.ne;
  const int npe = m.npe;
  std::vector<Index> inc_ptr((size_t)nn + 1, 0);
  for (Index e = 0; e < ne; ++e)
    for (int a = 0; a < npe; ++a) ++inc_ptr[(size_t)m.element_nodes(e)[a] + 1];
  for (Index n = 0; n < nn; ++n) inc_ptr[(size_t)n + 1] += inc_ptr[(size_t)n];
  std::vector<Index> inc((size_t)inc_ptr.back());
  {
    std::vector<Index> fill(inc_ptr.begin(), inc_ptr.end() - 1);
    for (Index e = 0; e < ne; ++e)
      for (int a = 0; a < npe; ++a) inc[(size_t)fill[(size_t)m.element_nodes(e)[a]]++] = e;
  }
  std::vector<Index> nbr_ptr((size_t)nn + 1, 0), nbr;
  {
    std::vector<Index> marker((size_t)nn, -1), list;
    for (Index b = 0; b < nn; ++b) {
      list.clear();
      for (Index k = inc_ptr[(size_t)b]; k < inc_ptr[(size_t)b + 1]; ++k) {
        const Index* nodes = m.element_nodes(inc[(size_t)k]);
        for (int a = 0; a < npe; ++a)
          if (marker[(size_t)nodes[a]] != b) {
            marker[(size_t)nodes[a]] = b;
            list.push_back(nodes[a]);
          }
      }
      std::sort(list.begin(), list.end());
      nbr.insert(nbr.end(), list.begin(), list.end());
      nbr_ptr[(size_t)b + 1] = (Index)nbr.size();
    }
  }
  Pattern p;
  p.n = nn * dim;
  p.outer.assign((size_t)p.n + 1, 0);
  for (Index b = 0; b < nn; ++b) {
    const Index deg = nbr_ptr[(size_t)b + 1] - nbr_ptr[(size_t)b];
    for (int kb = 0; kb < dim; ++kb) {
      const Index c = b * dim + kb;
      p.outer[(size_t)c + 1] = p.outer[(size_t)c] + deg * dim;
    }
  }
  p.inner.resize((size_t)p.outer.back());
  for (Index b = 0; b < nn; ++b)
    for (int kb = 0; kb < dim; ++kb) {
      Index q = p.outer[(size_t)(b * dim + kb)];
      for (Index k = nbr_ptr[(size_t)b]; k < nbr_ptr[(size_t)b + 1]; ++k)
        for (int ka = 0; ka < dim; ++ka) p.inner[(size_t)q++] = nbr[(size_t)k] * dim + ka;
    }
  p.block_offset.resize((size_t)ne * npe * npe);
  for (Index e = 0; e < ne; ++e) {
    const Index* nodes = m.element_nodes(e);
    for (int b = 0; b < npe; ++b) {
      const Index* begin = nbr.data() + nbr_ptr[(size_t)nodes[b]];
      const Index* end = nbr.data() + nbr_ptr[(size_t)nodes[b] + 1];
      for (int a = 0; a < npe; ++a) {
        const Index* hit = std::lower_bound(begin, end, nodes[a]);
        p.block_offset[((size_t)e * npe + a) * npe + b] = (Index)(hit - begin) * dim;
      }
    }
  }
  return p;
}

void scatter(const Pattern& p, const Mesh& m, int dim, const double* ke, const double* scale,
             std::vector<double>& values) {  // Assembler.cpp:182-214
  const int npe = m.npe;
  const int nd = npe * dim;
  values.assign(p.inner.size(), 0.0);
  for (Index e = 0; e < m.ne; ++e) {
    const double s = scale ? scale[e] : 1.0;
    const Index* nodes = m.element_nodes(e);
    const Index* offsets = p.block_offset.data() + (size_t)e * npe * npe;
    for (int b = 0; b < npe; ++b)
      for (int kb = 0; kb < dim; ++kb) {
        const int column = dim * b + kb;
        const Index base = p.outer[(size_t)(nodes[b] * dim + kb)];
        for (int a = 0; a < npe; ++a) {
          double* target = values.data() + base + offsets[a * npe + b];
          for (int ka = 0; ka < dim; ++ka) target[ka] += s * ke[(dim * a + ka) * nd + column];
        }
      }
  }
}

Reduction free_free_map(const Pattern& p, const Dofs& dofs) {  // Assembler.cpp:410-438
  Reduction r;
  const Index nf = (Index)dofs.free.size();
  r.rows = nf;
  r.ptr.assign((size_t)nf + 1, 0);
  for (Index rc = 0; rc < nf; ++rc) {
    const Index col = dofs.free[(size_t)rc];
    for (Index k = p.outer[(size_t)col]; k < p.outer[(size_t)col + 1]; ++k) {
      const Index rr = dofs.reduced[(size_t)p.inner[(size_t)k]];
      if (rr < 0) continue;
      r.col.push_back(rr);
      r.src.push_back(k);
    }
    r.ptr[(size_t)rc + 1] = (Index)r.col.size();
  }
  return r;
}

Csr reduce(const Reduction& r, const Pattern&, const std::vector<double>& full) {
  Csr c;
  c.rows = c.cols = r.rows;
  c.ptr = r.ptr;
  c.col = r.col;
  c.val.resize(r.src.size());
  for (size_t k = 0; k < r.src.size(); ++k) c.val[k] = full[(size_t)r.src[k]];
  return c;
}

// ---------------------------------------------------------------------------
// Sparse kernels (Multigrid.cpp:74-284)
// ---------------------------------------------------------------------------
double chunked_dot(const double* a, const double* b, Index n) {
  const Index chunks = (n + kChunk - 1) / kChunk;
  if (chunks <= 1) {
    double s = 0.0;
    for (Index i = 0; i < n; ++i) s += a[i] * b[i];
    return 0.0 + s;
  }
  std::vector<double> partial((size_t)chunks, 0.0);
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
  for (Index c = 0; c < chunks; ++c) {
    const Index begin = c * kChunk, end = std::min(n, begin + kChunk);
    double s = 0.0;
    for (Index i = begin; i < end; ++i) s += a[i] * b[i];
    partial[(size_t)c] = s;
  }
  double total = 0.0;
  for (double p : partial) total += p;
  return total;
}

void spmv(const Csr& a, const double* x, double* y) {
#pragma omp parallel for schedule(static) if (a.rows >= kParallelRows)
  for (Index i = 0; i < a.rows; ++i) {
    double s = 0.0;
    for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) s += a.val[(size_t)k] * x[a.col[(size_t)k]];
    y[i] = s;
  }
}

void residual(const Csr& a, const double* x, const double* b, double* r) {
#pragma omp parallel for schedule(static) if (a.rows >= kParallelRows)
  for (Index i = 0; i < a.rows; ++i) {
    double s = 0.0;
    for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) s += a.val[(size_t)k] * x[a.col[(size_t)k]];
    r[i] = b[i] - s;
  }
}

Csr multiply(const Csr& a, const Csr& b) {
  if (a.cols != b.rows) fail(2, "multigrid: inner dimensions of a product differ");
  Csr c;
  c.rows = a.rows;
  c.cols = b.cols;
  c.ptr.assign((size_t)a.rows + 1, 0);
#pragma omp parallel if (a.rows >= kParallelRows)
  {
    std::vector<Index> marker((size_t)b.cols, -1);
#pragma omp for schedule(static)
    for (Index i = 0; i < a.rows; ++i) {
      Index count = 0;
      for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
        const Index j = a.col[(size_t)k];
        for (Index kk = b.ptr[(size_t)j]; kk < b.ptr[(size_t)j + 1]; ++kk) {
          const Index col = b.col[(size_t)kk];
          if (marker[(size_t)col] != i) {
            marker[(size_t)col] = i;
            ++count;
          }
        }
      }
      c.ptr[(size_t)i + 1] = count;
    }
  }
  for (Index i = 0; i < a.rows; ++i) c.ptr[(size_t)i + 1] += c.ptr[(size_t)i];
  c.col.resize((size_t)c.nnz());
  c.val.resize((size_t)c.nnz());
#pragma omp parallel if (a.rows >= kParallelRows)
  {
    std::vector<Index> marker((size_t)b.cols, -1), cols;
    std::vector<double> acc((size_t)b.cols, 0.0);
#pragma omp for schedule(static)
    for (Index i = 0; i < a.rows; ++i) {
      cols.clear();
      for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
        const Index j = a.col[(size_t)k];
        const double av = a.val[(size_t)k];
        for (Index kk = b.ptr[(size_t)j]; kk < b.ptr[(size_t)j + 1]; ++kk) {
          const size_t sc = (size_t)b.col[(size_t)kk];
          if (marker[sc] != i) {
            marker[sc] = i;
            acc[sc] = av * b.val[(size_t)kk];
            cols.push_back(b.col[(size_t)kk]);
          } else {
            acc[sc] += av * b.val[(size_t)kk];
          }
        }
      }
      std::sort(cols.begin(), cols.end());
      Index p = c.ptr[(size_t)i];
      for (Index col : cols) {
        c.col[(size_t)p] = col;
        c.val[(size_t)p] = acc[(size_t)col];
        ++p;
      }
    }
  }
  return c;
}

void multiply_numeric(const Csr& a, const Csr& b, Csr& c) {
#pragma omp parallel if (a.rows >= kParallelRows)
  {
    std::vector<Index> slot((size_t)b.cols, 0);
#pragma omp for schedule(static)
    for (Index i = 0; i < a.rows; ++i) {
      for (Index p = c.ptr[(size_t)i]; p < c.ptr[(size_t)i + 1]; ++p) {
        slot[(size_t)c.col[(size_t)p]] = p;
        c.val[(size_t)p] = 0.0;
      }
      for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
        const Index j = a.col[(size_t)k];
        const double av = a.val[(size_t)k];
        for (Index kk = b.ptr[(size_t)j]; kk < b.ptr[(size_t)j + 1]; ++kk)
          c.val[(size_t)slot[(size_t)b.col[(size_t)kk]]] += av * b.val[(size_t)kk];
      }
    }
  }
}

Csr transpose(const Csr& a) {
  Csr t;
  t.rows = a.cols;
  t.cols = a.rows;
  t.ptr.assign((size_t)a.cols + 1, 0);
  for (Index k = 0; k < a.nnz(); ++k) ++t.ptr[(size_t)a.col[(size_t)k] + 1];
  for (Index j = 0; j < a.cols; ++j) t.ptr[(size_t)j + 1] += t.ptr[(size_t)j];
  t.col.resize((size_t)a.nnz());
  t.val.resize((size_t)a.nnz());
  std::vector<Index> next(t.ptr.begin(), t.ptr.end() - 1);
  for (Index i = 0; i < a.rows; ++i)
    for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
      const Index p = next[(size_t)a.col[(size_t)k]]++;
      t.col[(size_t)p] = i;
      t.val[(size_t)p] = a.val[(size_t)k];
    }
  return t;
}

Csr symmetrize(const Csr& a) {
  const Csr t = transpose(a);
  Csr s;
  s.rows = a.rows;
  s.cols = a.cols;
  s.ptr.assign((size_t)a.rows + 1, 0);
  s.col.reserve((size_t)a.nnz());
  s.val.reserve((size_t)a.nnz());
  const Index big = std::numeric_limits<Index>::max();
  for (Index i = 0; i < a.rows; ++i) {
    Index p = a.ptr[(size_t)i], q = t.ptr[(size_t)i];
    const Index pe = a.ptr[(size_t)i + 1], qe = t.ptr[(size_t)i + 1];
    while (p < pe || q < qe) {
      const Index cp = p < pe ? a.col[(size_t)p] : big;
      const Index cq = q < qe ? t.col[(size_t)q] : big;
      if (cp == cq) {
        s.col.push_back(cp);
        s.val.push_back(0.5 * (a.val[(size_t)p] + t.val[(size_t)q]));
        ++p;
        ++q;
      } else if (cp < cq) {
        s.col.push_back(cp);
        s.val.push_back(0.5 * a.val[(size_t)p]);
        ++p;
      } else {
        s.col.push_back(cq);
        s.val.push_back(0.5 * t.val[(size_t)q]);
        ++q;
      }
    }
    s.ptr[(size_t)i + 1] = (Index)s.col.size();
  }
  return s;
}

namespace {

// Plane rotation [c s; -s c] with [c s; -s c]^T [p; q] = [r; 0], in the sign
// convention of Eigen 3.4.0's JacobiRotation::makeGivens (real case).
void givens_rotation(double p, double q, double& c, double& s) {
  if (q == 0.0) {
    c = p < 0.0 ? -1.0 : 1.0;
    s = 0.0;
  } else if (p == 0.0) {
    c = 0.0;
    s = q < 0.0 ? 1.0 : -1.0;
  } else if (std::abs(p) > std::abs(q)) {
    const double t = q / p;
    double u = std::sqrt(1.0 + t * t);
    if (p < 0.0) u = -u;
    c = 1.0 / u;
    s = -t * c;
  } else {
    const double t = p / q;
    double u = std::sqrt(1.0 + t * t);
    if (q < 0.0) u = -u;
    s = -1.0 / u;
    c = -t * s;
  }
}

// sqrt(x^2 + y^2) for x, y >= 0 as max * sqrt(1 + (min / max)^2)
double scaled_hypot(double x, double y) {
  const double p = std::max(x, y);
  if (p == 0.0) return 0.0;
  const double r = std::min(y, x) / p;
  return p * std::sqrt(1.0 + r * r);
}

// One implicit symmetric QR step with the Wilkinson shift on the unreduced
// block [start, end] (Golub & Van Loan, Algorithm 8.3.2), chasing the bulge
// with plane rotations.
void tridiagonal_qr_step(double* d, double* e, int start, int end) {
  const double td = (d[end - 1] - d[end]) * 0.5;
  const double f = e[end - 1];
  double mu = d[end];
  if (td == 0.0) {
    mu -= std::abs(f);
  } else if (f != 0.0) {
    const double f2 = f * f;
    const double h = scaled_hypot(std::abs(td), std::abs(f));
    if (f2 == 0.0)
      mu -= f / ((td + (td > 0.0 ? h : -h)) / f);
    else
      mu -= f2 / (td + (td > 0.0 ? h : -h));
  }
  double x = d[start] - mu, z = e[start];
  for (int k = start; k < end && z != 0.0; ++k) {
    double c, s;
    givens_rotation(x, z, c, s);
    const double sdk = s * d[k] + c * e[k];
    const double dkp1 = s * e[k] + c * d[k + 1];
    d[k] = c * (c * d[k] - s * e[k]) - s * (c * e[k] - s * d[k + 1]);
    d[k + 1] = s * sdk + c * dkp1;
    e[k] = c * sdk - s * dkp1;
    if (k > start) e[k - 1] = c * e[k - 1] - s * z;
    x = e[k];
    if (k < end - 1) {
      z = -s * e[k + 1];
      e[k + 1] = c * e[k + 1];
    }
  }
}

}  // namespace

double tridiagonal_max_eigenvalue(const std::vector<double>& diag, const std::vector<double>& sub) {
  // Multigrid.cpp:324-326 takes the largest eigenvalue from Eigen 3.4.0's
  // SelfAdjointEigenSolver::computeFromTridiagonal (eigenvalues only, no
  // scaling, at most 30 n sweeps). That routine is the implicit symmetric QR
  // iteration with Wilkinson shifts; restated here with its deflation rule: an
  // off-diagonal entry is set to zero when it is below the smallest normal
  // number or when (e_i / eps)^2 <= |d_i| + |d_(i+1)|; the trailing unreduced
  // block is then stepped.
  const int k = (int)diag.size();
  if (k == 1) return diag[0];
  std::vector<double> d(diag), e(sub.begin(), sub.begin() + (k - 1));
  const double tiny = std::numeric_limits<double>::min();
  const double inv_eps = 1.0 / std::numeric_limits<double>::epsilon();
  int end = k - 1, start = 0;
  long long iter = 0;
  while (end > 0) {
    for (int i = start; i < end; ++i) {
      if (std::abs(e[(size_t)i]) < tiny) {
        e[(size_t)i] = 0.0;
      } else {
        const double se = inv_eps * e[(size_t)i];
        if (se * se <= std::abs(d[(size_t)i]) + std::abs(d[(size_t)i + 1])) e[(size_t)i] = 0.0;
      }
    }
    while (end > 0 && e[(size_t)end - 1] == 0.0) --end;
    if (end <= 0) break;
    if (++iter > 30LL * k) break;
    start = end - 1;
    while (start > 0 && e[(size_t)start - 1] != 0.0) --start;
    tridiagonal_qr_step(d.data(), e.data(), start, end);
  }
  double best = d[0];
  for (int i = 1; i < k; ++i) best = std::max(best, d[(size_t)i]);
  return best;
}

double lanczos_lambda_max(const Csr& a, const std::vector<double>& inv_diag, int steps) {
  const Index n = a.rows;  // Multigrid.cpp:288-327
  if (n == 0) return 0.0;
  std::vector<double> s((size_t)n), v((size_t)n), v_prev((size_t)n, 0.0), w((size_t)n), tmp((size_t)n);
  for (Index i = 0; i < n; ++i) s[(size_t)i] = std::sqrt(inv_diag[(size_t)i]);
  for (Index i = 0; i < n; ++i) v[(size_t)i] = 1.0 + 0.5 * std::sin(1.7 * double(i) + 0.3);
  const double vn = std::sqrt(chunked_dot(v.data(), v.data(), n));
  for (auto& t : v) t /= vn;
  std::vector<double> alpha, beta;
  const int m = std::max(1, std::min<int>(steps, (int)n));
  for (int j = 0; j < m; ++j) {
    for (Index i = 0; i < n; ++i) tmp[(size_t)i] = s[(size_t)i] * v[(size_t)i];
    spmv(a, tmp.data(), w.data());
    for (Index i = 0; i < n; ++i) w[(size_t)i] *= s[(size_t)i];
    if (j > 0) {
      const double bb = beta.back();
      for (Index i = 0; i < n; ++i) w[(size_t)i] -= bb * v_prev[(size_t)i];
    }
    const double aj = chunked_dot(w.data(), v.data(), n);
    alpha.push_back(aj);
    for (Index i = 0; i < n; ++i) w[(size_t)i] -= aj * v[(size_t)i];
    const double bj = std::sqrt(chunked_dot(w.data(), w.data(), n));
    if (j + 1 == m || !(bj > 1.0e-12 * std::abs(aj))) break;
    beta.push_back(bj);
    v_prev = v;
    for (Index i = 0; i < n; ++i) v[(size_t)i] = w[(size_t)i] / bj;
  }
  std::vector<double> sub(beta.begin(), beta.begin() + (alpha.size() - 1));
  return tridiagonal_max_eigenvalue(alpha, sub);
}

namespace {

// Reduction orders of Eigen 3.4.0 built for x86-64 with SSE2 and no -march
// flag (the reference's build), where a double packet holds two lanes.
//
// sum / squaredNorm / dot of a dynamic vector (redux, linear vectorised
// traversal, no unrolling): lanes {0,1} and {2,3} start two packet
// accumulators, each later step of four adds {i,i+1} to the first and
// {i+2,i+3} to the second; the two packets are added, then one more packet if
// two entries of the even part remain, then the packet's two lanes, then the
// odd last entry.
template <class F>
double packet_redux(Index n, F&& x) {
  if (n <= 0) return 0.0;
  const Index n2 = (n / 2) * 2;
  if (n2 == 0) return x(0);
  double p0a = x(0), p0b = x(1);
  if (n2 > 2) {
    double p1a = x(2), p1b = x(3);
    const Index n4 = (n / 4) * 4;
    for (Index i = 4; i < n4; i += 4) {
      p0a = p0a + x(i);
      p0b = p0b + x(i + 1);
      p1a = p1a + x(i + 2);
      p1b = p1b + x(i + 3);
    }
    p0a = p0a + p1a;
    p0b = p0b + p1b;
    if (n2 > n4) {
      p0a = p0a + x(n4);
      p0b = p0b + x(n4 + 1);
    }
  }
  double res = p0a + p0b;
  for (Index i = n2; i < n; ++i) res = res + x(i);
  return res;
}

// One output of the row-major matrix-vector kernel (row vector times a
// column-major block): even and odd entries accumulate in the two lanes of one
// packet (product, then add), the lanes are added, then the odd last product.
double packet_gemv_dot(const double* a, const double* b, Index n) {
  double c0 = 0.0, c1 = 0.0;
  Index j = 0;
  for (; j + 2 <= n; j += 2) {
    c0 = a[j] * b[j] + c0;
    c1 = a[j + 1] * b[j + 1] + c1;
  }
  double cc = c0 + c1;
  for (; j < n; ++j) cc = cc + a[j] * b[j];
  return 0.0 + cc;
}

// Applies H = I - tau [1; v] [1; v]^T to the rows x cols column-major block at
// blk (leading dimension ld) as MatrixBase::applyHouseholderOnTheLeft does:
// tmp = v^T bottom (the matrix-vector kernel, or a dot for one column),
// tmp += row 0, row 0 -= tau tmp, bottom -= (tau v) tmp.
void apply_householder_left(double* blk, Index ld, Index rows, Index cols, const double* v, double tau,
                            std::vector<double>& tmp, std::vector<double>& tv) {
  if (rows == 1) {
    const double f = 1.0 - tau;
    for (Index j = 0; j < cols; ++j) blk[(size_t)j * ld] *= f;
    return;
  }
  if (tau == 0.0 || cols == 0) return;
  const Index nb = rows - 1;
  tmp.assign((size_t)cols, 0.0);
  if (cols == 1) {
    const double* b = blk + 1;
    tmp[0] = 0.0 + packet_redux(nb, [&](Index i) { return v[i] * b[i]; });
  } else {
    for (Index j = 0; j < cols; ++j) tmp[(size_t)j] = packet_gemv_dot(blk + (size_t)j * ld + 1, v, nb);
  }
  for (Index j = 0; j < cols; ++j) tmp[(size_t)j] = tmp[(size_t)j] + blk[(size_t)j * ld];
  for (Index j = 0; j < cols; ++j) blk[(size_t)j * ld] = blk[(size_t)j * ld] - tau * tmp[(size_t)j];
  tv.resize((size_t)nb);
  for (Index i = 0; i < nb; ++i) tv[(size_t)i] = tau * v[i];
  for (Index j = 0; j < cols; ++j) {
    double* col = blk + (size_t)j * ld + 1;
    const double t = tmp[(size_t)j];
    for (Index i = 0; i < nb; ++i) col[i] = col[i] - t * tv[(size_t)i];
  }
}

}  // namespace

Index cpqr_block(const std::vector<double>& block, Index rows, Index m, std::vector<double>& q,
                 std::vector<double>& r) {
  // Multigrid.cpp:606-617: Eigen::ColPivHouseholderQR of the aggregate's
  // near-null-space rows, rank at 1e-10 |R00|, Q = householderQ() * I(rows,
  // rank), R = R(0:rank, :) P^T. Restated from Eigen 3.4.0's
  // ColPivHouseholderQR::computeInPlace (Businger-Golub column pivoting on the
  // updated norms, first largest wins; LAPACK xGEQP3 norm downdating with the
  // direct recomputation below sqrt(eps); Householder reflector beta =
  // -sign(c0) sqrt(c0^2 + |tail|^2), v = tail / (c0 - beta), tau = (beta - c0)
  // / beta, skipped when |tail|^2 <= DBL_MIN) and HouseholderSequence's
  // unblocked application (reflectors last to first on the identity), with
  // every sum in the order Eigen's SSE2 kernels use (packet_redux,
  // packet_gemv_dot above).
  std::vector<double> a((size_t)rows * m);  // column-major, like Eigen's storage
  for (Index i = 0; i < rows; ++i)
    for (Index j = 0; j < m; ++j) a[(size_t)j * rows + i] = block[(size_t)i * m + j];
  auto at = [&](Index i, Index j) -> double& { return a[(size_t)j * rows + i]; };
  auto norm_from = [&](Index j, Index from) {
    const double* c = a.data() + (size_t)j * rows + from;
    return std::sqrt(packet_redux(rows - from, [&](Index i) { return c[i] * c[i]; }));
  };
  std::vector<Index> perm((size_t)m);
  std::iota(perm.begin(), perm.end(), 0);
  const Index size = std::min(rows, m);
  std::vector<double> tau((size_t)size, 0.0), norm_upd((size_t)m), norm_dir((size_t)m), tmp, tv;
  for (Index j = 0; j < m; ++j) norm_dir[(size_t)j] = norm_upd[(size_t)j] = norm_from(j, 0);
  const double downdate_threshold = std::sqrt(std::numeric_limits<double>::epsilon());
  for (Index k = 0; k < size; ++k) {
    Index best = k;
    for (Index j = k + 1; j < m; ++j)
      if (norm_upd[(size_t)j] > norm_upd[(size_t)best]) best = j;
    if (best != k) {
      for (Index i = 0; i < rows; ++i) std::swap(at(i, k), at(i, best));
      std::swap(norm_upd[(size_t)k], norm_upd[(size_t)best]);
      std::swap(norm_dir[(size_t)k], norm_dir[(size_t)best]);
      std::swap(perm[(size_t)k], perm[(size_t)best]);
    }
    double* col = a.data() + (size_t)k * rows;
    const double c0 = col[k];
    const Index nt = rows - k - 1;
    const double tail_sq = nt > 0 ? packet_redux(nt, [&](Index i) { return col[k + 1 + i] * col[k + 1 + i]; }) : 0.0;
    double beta, t;
    if (tail_sq <= std::numeric_limits<double>::min()) {
      t = 0.0;
      beta = c0;
      for (Index i = k + 1; i < rows; ++i) col[i] = 0.0;
    } else {
      beta = std::sqrt(c0 * c0 + tail_sq);
      if (c0 >= 0.0) beta = -beta;
      const double den = c0 - beta;
      for (Index i = k + 1; i < rows; ++i) col[i] = col[i] / den;
      t = (beta - c0) / beta;
    }
    tau[(size_t)k] = t;
    col[k] = beta;
    apply_householder_left(a.data() + (size_t)(k + 1) * rows + k, rows, rows - k, m - k - 1, col + k + 1, t, tmp, tv);
    for (Index j = k + 1; j < m; ++j) {
      if (norm_upd[(size_t)j] == 0.0) continue;
      double temp = std::abs(at(k, j)) / norm_upd[(size_t)j];
      temp = (1.0 + temp) * (1.0 - temp);
      temp = temp < 0.0 ? 0.0 : temp;
      const double ratio = norm_upd[(size_t)j] / norm_dir[(size_t)j];
      const double temp2 = temp * (ratio * ratio);
      if (temp2 <= downdate_threshold) {
        norm_dir[(size_t)j] = norm_from(j, k + 1);
        norm_upd[(size_t)j] = norm_dir[(size_t)j];
      } else {
        norm_upd[(size_t)j] *= std::sqrt(temp);
      }
    }
  }
  const Index kmax = std::min(rows, m);
  const double lead = kmax > 0 ? std::abs(at(0, 0)) : 0.0;
  Index rank = 0;
  for (Index i = 0; i < kmax; ++i)
    if (std::abs(at(i, i)) > 1.0e-10 * lead) ++rank;
  rank = std::max<Index>(rank, 1);
  // Q(:, 0:rank): the reflectors from the last to the first, each on rows k..rows-1
  std::vector<double> qc((size_t)rows * rank, 0.0);  // column-major
  for (Index c = 0; c < rank; ++c) qc[(size_t)c * rows + c] = 1.0;
  for (Index k = size - 1; k >= 0; --k)
    apply_householder_left(qc.data() + k, rows, rows - k, rank, a.data() + (size_t)k * rows + k + 1, tau[(size_t)k], tmp,
                           tv);
  q.assign((size_t)rows * rank, 0.0);
  for (Index i = 0; i < rows; ++i)
    for (Index c = 0; c < rank; ++c) q[(size_t)i * rank + c] = qc[(size_t)c * rows + i];
  // R(0:rank, :) un-pivoted: column perm[j] of the result holds column j of R.
  r.assign((size_t)rank * m, 0.0);
  for (Index i = 0; i < rank; ++i)
    for (Index j = i; j < m; ++j) r[(size_t)i * m + perm[(size_t)j]] = at(i, j);
  return rank;
}

// ---------------------------------------------------------------------------
// SA-AMG (Multigrid.cpp:335-1111)
// ---------------------------------------------------------------------------
Amg::Amg(AmgOptions o) : opt_(o) {  // Multigrid.cpp:908-925
  if (!(o.strength_threshold >= 0.0 && o.strength_threshold < 1.0)) fail(1, "multigrid strength_threshold must lie in [0, 1)");
  if (o.max_levels < 1) fail(1, "multigrid max_levels must be at least 1");
  if (o.coarse_size < 1) fail(1, "multigrid coarse_size must be positive");
  if (o.smoother_degree < 1) fail(1, "multigrid smoother_degree must be at least 1");
  if (!(o.chebyshev_ratio > 1.0)) fail(1, "multigrid chebyshev_ratio must exceed 1");
  if (!(o.prolongator_damping > 0.0 && o.prolongator_damping < 2.0)) fail(1, "multigrid prolongator_damping must lie in (0, 2)");
  if (o.lanczos_steps < 2) fail(1, "multigrid lanczos_steps must be at least 2");
}

void Amg::build_finest_nodes(const Layout& layout, Index n) {
  Level& lvl = levels_.front();
  const int dim = layout.dim, ndpn = layout.ndpn > 0 ? layout.ndpn : layout.dim;
  lvl.node_ptr.assign(1, 0);
  Index previous = -1;
  for (Index i = 0; i < n; ++i) {
    const Index g = layout.unknowns[i];
    const Index node = g / ndpn;
    if (node >= layout.nn || g < 0) fail(2, "multigrid: an unknown maps outside the model's nodes");
    if (i > 0 && g <= layout.unknowns[i - 1]) fail(2, "multigrid: the DOF layout must list unknowns in ascending order");
    if (node != previous) {
      if (i > 0) lvl.node_ptr.push_back(i);
      previous = node;
    }
  }
  lvl.node_ptr.push_back(n);
  lvl.num_nodes = (Index)lvl.node_ptr.size() - 1;
  double lo[3], hi[3];
  for (int d = 0; d < 3; ++d) {
    lo[d] = std::numeric_limits<double>::max();
    hi[d] = -std::numeric_limits<double>::max();
  }
  for (Index c = 0; c < layout.nn; ++c)
    for (int d = 0; d < 3; ++d) {
      const double x = d < dim ? layout.coords[(size_t)c * dim + d] : 0.0;
      lo[d] = std::min(lo[d], x);
      hi[d] = std::max(hi[d], x);
    }
  double centre[3], scale = 0.0;
  for (int d = 0; d < 3; ++d) {
    centre[d] = 0.5 * (lo[d] + hi[d]);
    scale = std::max(scale, hi[d] - lo[d]);
  }
  scale = std::max(scale, 1.0e-300);
  m_ = dim == 2 ? 3 : 6;
  lvl.nullspace.assign((size_t)n * m_, 0.0);
  for (Index i = 0; i < n; ++i) {
    const Index g = layout.unknowns[i];
    const Index node = g / ndpn;
    const int k = (int)(g % ndpn);
    double r[3];
    for (int d = 0; d < 3; ++d) r[d] = ((d < dim ? layout.coords[(size_t)node * dim + d] : 0.0) - centre[d]) / scale;
    double* row = &lvl.nullspace[(size_t)i * m_];
    if (k >= 3) {
      row[k] = 1.0 / scale;
      continue;
    }
    row[k] = 1.0;
    if (dim == 2) {
      row[2] = k == 0 ? -r[1] : r[0];
    } else {
      // u = omega x r for unit rotations about x, y and z
      const double rx[3] = {0.0, -r[2], r[1]};
      const double ry[3] = {r[2], 0.0, -r[0]};
      const double rz[3] = {-r[1], r[0], 0.0};
      row[3] = rx[k];
      row[4] = ry[k];
      row[5] = rz[k];
    }
  }
  stats_.near_null_space_dimension = m_;
}

void Amg::aggregate_level(Level& lvl) {
  const Csr& a = lvl.a;
  const Index nn = lvl.num_nodes;
  std::vector<Index> node_of((size_t)a.rows);
  for (Index node = 0; node < nn; ++node)
    for (Index i = lvl.node_ptr[(size_t)node]; i < lvl.node_ptr[(size_t)node + 1]; ++i) node_of[(size_t)i] = node;
  std::vector<double> block_diag((size_t)nn, 0.0);
  for (Index node = 0; node < nn; ++node) {
    double s = 0.0;
    for (Index i = lvl.node_ptr[(size_t)node]; i < lvl.node_ptr[(size_t)node + 1]; ++i)
      for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k)
        if (node_of[(size_t)a.col[(size_t)k]] == node) s += a.val[(size_t)k] * a.val[(size_t)k];
    block_diag[(size_t)node] = std::sqrt(s);
  }
  std::vector<Index> strong_ptr((size_t)nn + 1, 0), strong;
  std::vector<double> strong_value;
  {
    std::vector<Index> marker((size_t)nn, -1), touched;
    std::vector<double> acc((size_t)nn, 0.0);
    const double theta = opt_.strength_threshold;
    for (Index node = 0; node < nn; ++node) {
      touched.clear();
      for (Index i = lvl.node_ptr[(size_t)node]; i < lvl.node_ptr[(size_t)node + 1]; ++i)
        for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
          const Index other = node_of[(size_t)a.col[(size_t)k]];
          if (other == node) continue;
          if (marker[(size_t)other] != node) {
            marker[(size_t)other] = node;
            acc[(size_t)other] = 0.0;
            touched.push_back(other);
          }
          acc[(size_t)other] += a.val[(size_t)k] * a.val[(size_t)k];
        }
      std::sort(touched.begin(), touched.end());
      const double di = block_diag[(size_t)node];
      for (Index other : touched) {
        const double s = std::sqrt(acc[(size_t)other]);
        const double ref = std::sqrt(di * block_diag[(size_t)other]);
        if (s > 0.0 && s >= theta * ref) {
          strong.push_back(other);
          strong_value.push_back(ref > 0.0 ? s / ref : 0.0);
        }
      }
      strong_ptr[(size_t)node + 1] = (Index)strong.size();
    }
  }
  // Greedy aggregation, three passes in node order (Multigrid.cpp:518-564)
  std::vector<Index>& agg = lvl.aggregate;
  agg.assign((size_t)nn, -1);
  Index count = 0;
  for (Index node = 0; node < nn; ++node) {
    if (agg[(size_t)node] >= 0) continue;
    const Index b = strong_ptr[(size_t)node], e = strong_ptr[(size_t)node + 1];
    if (b == e) continue;
    bool free_nb = true;
    for (Index k = b; k < e; ++k)
      if (agg[(size_t)strong[(size_t)k]] >= 0) {
        free_nb = false;
        break;
      }
    if (!free_nb) continue;
    agg[(size_t)node] = count;
    for (Index k = b; k < e; ++k) agg[(size_t)strong[(size_t)k]] = count;
    ++count;
  }
  const std::vector<Index> first = agg;
  for (Index node = 0; node < nn; ++node) {
    if (first[(size_t)node] >= 0) continue;
    Index best = -1;
    double best_value = -1.0;
    for (Index k = strong_ptr[(size_t)node]; k < strong_ptr[(size_t)node + 1]; ++k) {
      const Index target = first[(size_t)strong[(size_t)k]];
      if (target >= 0 && strong_value[(size_t)k] > best_value) {
        best = target;
        best_value = strong_value[(size_t)k];
      }
    }
    if (best >= 0) agg[(size_t)node] = best;
  }
  for (Index node = 0; node < nn; ++node) {
    if (agg[(size_t)node] >= 0) continue;
    agg[(size_t)node] = count;
    for (Index k = strong_ptr[(size_t)node]; k < strong_ptr[(size_t)node + 1]; ++k) {
      const Index other = strong[(size_t)k];
      if (agg[(size_t)other] < 0) agg[(size_t)other] = count;
    }
    ++count;
  }
}

void Amg::tentative_prolongator(Level& lvl, Level& next) {
  const Index n = lvl.a.rows, nn = lvl.num_nodes, m = m_;
  const std::vector<Index>& agg = lvl.aggregate;
  const Index na = agg.empty() ? 0 : *std::max_element(agg.begin(), agg.end()) + 1;
  std::vector<Index> agg_ptr((size_t)na + 1, 0), agg_nodes((size_t)nn);
  for (Index node = 0; node < nn; ++node) ++agg_ptr[(size_t)agg[(size_t)node] + 1];
  for (Index a = 0; a < na; ++a) agg_ptr[(size_t)a + 1] += agg_ptr[(size_t)a];
  {
    std::vector<Index> fill(agg_ptr.begin(), agg_ptr.end() - 1);
    for (Index node = 0; node < nn; ++node) agg_nodes[(size_t)fill[(size_t)agg[(size_t)node]]++] = node;
  }
  std::vector<std::vector<double>> q((size_t)na), r((size_t)na);
  std::vector<Index> rank((size_t)na);
#pragma omp parallel for schedule(dynamic, 256) if (na >= 4096)
  for (Index a = 0; a < na; ++a) {
    Index rows = 0;
    for (Index k = agg_ptr[(size_t)a]; k < agg_ptr[(size_t)a + 1]; ++k) {
      const Index node = agg_nodes[(size_t)k];
      rows += lvl.node_ptr[(size_t)node + 1] - lvl.node_ptr[(size_t)node];
    }
    std::vector<double> block((size_t)rows * m);
    Index row = 0;
    for (Index k = agg_ptr[(size_t)a]; k < agg_ptr[(size_t)a + 1]; ++k) {
      const Index node = agg_nodes[(size_t)k];
      for (Index i = lvl.node_ptr[(size_t)node]; i < lvl.node_ptr[(size_t)node + 1]; ++i, ++row)
        std::copy(&lvl.nullspace[(size_t)i * m], &lvl.nullspace[(size_t)i * m] + m, &block[(size_t)row * m]);
    }
    rank[(size_t)a] = cpqr_block(block, rows, m, q[(size_t)a], r[(size_t)a]);
  }
  next.node_ptr.assign((size_t)na + 1, 0);
  for (Index a = 0; a < na; ++a) next.node_ptr[(size_t)a + 1] = next.node_ptr[(size_t)a] + rank[(size_t)a];
  next.num_nodes = na;
  const Index nc = next.node_ptr.back();
  next.nullspace.assign((size_t)nc * m, 0.0);
  for (Index a = 0; a < na; ++a)
    std::copy(r[(size_t)a].begin(), r[(size_t)a].end(), &next.nullspace[(size_t)next.node_ptr[(size_t)a] * m]);
  Csr& pt = lvl.p_tent;
  pt.rows = n;
  pt.cols = nc;
  pt.ptr.assign((size_t)n + 1, 0);
  std::vector<Index> agg_of_row((size_t)n), local_row((size_t)n);
  for (Index a = 0; a < na; ++a) {
    Index local = 0;
    for (Index k = agg_ptr[(size_t)a]; k < agg_ptr[(size_t)a + 1]; ++k) {
      const Index node = agg_nodes[(size_t)k];
      for (Index i = lvl.node_ptr[(size_t)node]; i < lvl.node_ptr[(size_t)node + 1]; ++i) {
        agg_of_row[(size_t)i] = a;
        local_row[(size_t)i] = local++;
      }
    }
  }
  for (Index i = 0; i < n; ++i) pt.ptr[(size_t)i + 1] = pt.ptr[(size_t)i] + rank[(size_t)agg_of_row[(size_t)i]];
  pt.col.resize((size_t)pt.nnz());
  pt.val.resize((size_t)pt.nnz());
  for (Index i = 0; i < n; ++i) {
    const Index a = agg_of_row[(size_t)i];
    const Index ra = rank[(size_t)a];
    Index p = pt.ptr[(size_t)i];
    for (Index j = 0; j < ra; ++j, ++p) {
      pt.col[(size_t)p] = next.node_ptr[(size_t)a] + j;
      pt.val[(size_t)p] = q[(size_t)a][(size_t)local_row[(size_t)i] * ra + j];
    }
  }
}

void Amg::smooth_prolongator(Level& lvl, bool numeric_only) {
  const double omega = opt_.prolongator_damping / lvl.lambda_max;
  Csr ap;
  if (numeric_only) {
    ap = std::move(lvl.p);
    multiply_numeric(lvl.a, lvl.p_tent, ap);
  } else {
    ap = multiply(lvl.a, lvl.p_tent);
  }
  const Csr& pt = lvl.p_tent;
  bool lost = false;
#pragma omp parallel for schedule(static) if (ap.rows >= kParallelRows)
  for (Index i = 0; i < ap.rows; ++i) {
    const double scale = omega * lvl.inv_diag[(size_t)i];
    for (Index k = ap.ptr[(size_t)i]; k < ap.ptr[(size_t)i + 1]; ++k) ap.val[(size_t)k] *= -scale;
    for (Index kk = pt.ptr[(size_t)i]; kk < pt.ptr[(size_t)i + 1]; ++kk) {
      const Index c = pt.col[(size_t)kk];
      const Index* begin = ap.col.data() + ap.ptr[(size_t)i];
      const Index* end = ap.col.data() + ap.ptr[(size_t)i + 1];
      const Index* hit = std::lower_bound(begin, end, c);
      if (hit == end || *hit != c) {
        lost = true;
        continue;
      }
      ap.val[(size_t)(hit - ap.col.data())] += pt.val[(size_t)kk];
    }
  }
  if (lost) fail(2, "multigrid: internal error, the smoothed prolongator lost an entry");
  lvl.p = std::move(ap);
  if (numeric_only) {
    Csr& r = lvl.r;
    std::vector<Index> next(r.ptr.begin(), r.ptr.end() - 1);
    for (Index i = 0; i < lvl.p.rows; ++i)
      for (Index k = lvl.p.ptr[(size_t)i]; k < lvl.p.ptr[(size_t)i + 1]; ++k)
        r.val[(size_t)next[(size_t)lvl.p.col[(size_t)k]]++] = lvl.p.val[(size_t)k];
  } else {
    lvl.r = transpose(lvl.p);
  }
}

void Amg::numeric_level(size_t l) {
  Level& lvl = levels_[l];
  const Csr& a = lvl.a;
  lvl.inv_diag.assign((size_t)a.rows, 0.0);
  for (Index i = 0; i < a.rows; ++i) {
    double d = 0.0;
    for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k)
      if (a.col[(size_t)k] == i) d += a.val[(size_t)k];
    if (!(d > 0.0)) {
      std::ostringstream os;
      os << "multigrid level " << l << ": diagonal entry " << i << " is " << d
         << "; a stiffness matrix must have a positive diagonal";
      fail(2, os.str());
    }
    lvl.inv_diag[(size_t)i] = 1.0 / d;
  }
  lvl.lambda_max = lanczos_lambda_max(a, lvl.inv_diag, opt_.lanczos_steps);
  if (!(lvl.lambda_max > 0.0) || !std::isfinite(lvl.lambda_max))
    fail(2, "multigrid: the spectral estimate of D^-1 A is not positive; the matrix is not positive definite");
  lvl.x.assign((size_t)a.rows, 0.0);
  lvl.b.assign((size_t)a.rows, 0.0);
  lvl.res.assign((size_t)a.rows, 0.0);
  lvl.d.assign((size_t)a.rows, 0.0);
}

void Amg::coarse_factorize(Level& lvl) {
  // Dense LDL^T with diagonal pivoting (the coarsest level is at most a few
  // thousand unknowns); the pivot ratio check follows Multigrid.cpp:766-781.
  const Index n = lvl.a.rows;
  if (n > 4000) fail(2, "refcpu: the coarsest level is larger than the dense limit (4000)");
  ldl_.assign((size_t)n * n, 0.0);
  for (Index i = 0; i < n; ++i)
    for (Index k = lvl.a.ptr[(size_t)i]; k < lvl.a.ptr[(size_t)i + 1]; ++k) ldl_[(size_t)i * n + lvl.a.col[(size_t)k]] = lvl.a.val[(size_t)k];
  perm_.resize((size_t)n);
  std::iota(perm_.begin(), perm_.end(), 0);
  double* A = ldl_.data();
  std::vector<double> dk((size_t)n);
  for (Index k = 0; k < n; ++k) {
    // pivot: the largest magnitude remaining diagonal entry
    Index p = k;
    for (Index i = k + 1; i < n; ++i)
      if (std::abs(A[(size_t)i * n + i]) > std::abs(A[(size_t)p * n + p])) p = i;
    if (p != k) {
      for (Index j = 0; j < n; ++j) std::swap(A[(size_t)k * n + j], A[(size_t)p * n + j]);
      for (Index i = 0; i < n; ++i) std::swap(A[(size_t)i * n + k], A[(size_t)i * n + p]);
      std::swap(perm_[(size_t)k], perm_[(size_t)p]);
    }
    const double d = A[(size_t)k * n + k];
    dk[(size_t)k] = d;
    if (d == 0.0) continue;
    for (Index i = k + 1; i < n; ++i) {
      const double lik = A[(size_t)i * n + k] / d;
      const double* rowk = A + (size_t)k * n;
      double* rowi = A + (size_t)i * n;
      for (Index j = k + 1; j <= i; ++j) rowi[j] -= lik * rowk[j];
      rowi[k] = lik;
    }
    for (Index i = k + 1; i < n; ++i)  // keep the upper triangle symmetric for pivoting
      for (Index j = i + 1; j < n; ++j) A[(size_t)i * n + j] = A[(size_t)j * n + i];
  }
  double max_abs = 0.0, min_d = std::numeric_limits<double>::max();
  for (double d : dk) {
    max_abs = std::max(max_abs, std::abs(d));
    min_d = std::min(min_d, d);
  }
  const double ratio = max_abs > 0.0 ? min_d / max_abs : 0.0;
  stats_.coarse_pivot_ratio = ratio;
  if (!(ratio > opt_.coarse_pivot_tolerance)) {
    std::ostringstream os;
    os << "multigrid: the coarsest operator (" << n << " unknowns) is singular or indefinite (smallest / largest pivot "
       << ratio << "). A rigid-body mode survived to the coarse level, so the model is under-constrained";
    fail(2, os.str());
  }
}

void Amg::smooth(const Level& lvl, bool pre, bool zero_guess) const {
  const Csr& a = lvl.a;
  const Index n = a.rows;
  std::vector<double>& x = lvl.x;
  const std::vector<double>& b = lvl.b;
  if (opt_.smoother == 1) {  // Multigrid.cpp:788-806
    if (zero_guess) std::fill(x.begin(), x.end(), 0.0);
    for (int sweep = 0; sweep < opt_.smoother_degree; ++sweep) {
      if (pre) {
        for (Index i = 0; i < n; ++i) {
          double s = b[(size_t)i];
          for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) s -= a.val[(size_t)k] * x[(size_t)a.col[(size_t)k]];
          x[(size_t)i] += s * lvl.inv_diag[(size_t)i];
        }
      } else {
        for (Index i = n - 1; i >= 0; --i) {
          double s = b[(size_t)i];
          for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) s -= a.val[(size_t)k] * x[(size_t)a.col[(size_t)k]];
          x[(size_t)i] += s * lvl.inv_diag[(size_t)i];
        }
      }
    }
    return;
  }
  // Chebyshev (Saad, Algorithm 12.1) on [1.1 lmax / ratio, 1.1 lmax]
  const double upper = 1.1 * lvl.lambda_max;
  const double lower = upper / opt_.chebyshev_ratio;
  const double theta = 0.5 * (upper + lower);
  const double delta = 0.5 * (upper - lower);
  const double sigma = theta / delta;
  double rho = 1.0 / sigma;
  std::vector<double>& r = lvl.res;
  std::vector<double>& d = lvl.d;
  if (zero_guess) {
    std::fill(x.begin(), x.end(), 0.0);
    r = b;
  } else {
    residual(a, x.data(), b.data(), r.data());
  }
  const std::vector<double>& dinv = lvl.inv_diag;
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
  for (Index i = 0; i < n; ++i) {
    d[(size_t)i] = dinv[(size_t)i] * r[(size_t)i] / theta;
    x[(size_t)i] += d[(size_t)i];
  }
  for (int k = 1; k < opt_.smoother_degree; ++k) {
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
    for (Index i = 0; i < n; ++i) {
      double s = 0.0;
      for (Index kk = a.ptr[(size_t)i]; kk < a.ptr[(size_t)i + 1]; ++kk) s += a.val[(size_t)kk] * d[(size_t)a.col[(size_t)kk]];
      r[(size_t)i] -= s;
    }
    const double rho_new = 1.0 / (2.0 * sigma - rho);
    const double c1 = rho_new * rho;
    const double c2 = 2.0 * rho_new / delta;
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
    for (Index i = 0; i < n; ++i) {
      d[(size_t)i] = c1 * d[(size_t)i] + c2 * dinv[(size_t)i] * r[(size_t)i];
      x[(size_t)i] += d[(size_t)i];
    }
    rho = rho_new;
  }
}

void Amg::cycle(size_t l) const {
  const Level& lvl = levels_[l];
  if (l + 1 == levels_.size()) {
    const Index n = lvl.a.rows;
    std::vector<double> y((size_t)n);
    for (Index i = 0; i < n; ++i) y[(size_t)i] = lvl.b[(size_t)perm_[(size_t)i]];
    const double* A = ldl_.data();
    for (Index i = 0; i < n; ++i) {  // L y = P b
      double s = y[(size_t)i];
      for (Index j = 0; j < i; ++j) s -= A[(size_t)i * n + j] * y[(size_t)j];
      y[(size_t)i] = s;
    }
    for (Index i = 0; i < n; ++i) y[(size_t)i] /= A[(size_t)i * n + i];
    for (Index i = n - 1; i >= 0; --i) {  // L^T z = y
      double s = y[(size_t)i];
      for (Index j = i + 1; j < n; ++j) s -= A[(size_t)j * n + i] * y[(size_t)j];
      y[(size_t)i] = s;
    }
    for (Index i = 0; i < n; ++i) lvl.x[(size_t)perm_[(size_t)i]] = y[(size_t)i];
    return;
  }
  smooth(lvl, true, true);
  residual(lvl.a, lvl.x.data(), lvl.b.data(), lvl.res.data());
  const Level& next = levels_[l + 1];
  spmv(lvl.r, lvl.res.data(), next.b.data());
  cycle(l + 1);
  const Csr& p = lvl.p;
#pragma omp parallel for schedule(static) if (p.rows >= kParallelRows)
  for (Index i = 0; i < p.rows; ++i) {
    double s = 0.0;
    for (Index k = p.ptr[(size_t)i]; k < p.ptr[(size_t)i + 1]; ++k) s += p.val[(size_t)k] * next.x[(size_t)p.col[(size_t)k]];
    lvl.x[(size_t)i] += s;
  }
  smooth(lvl, false, false);
}

void Amg::setup(const Csr& a, const Layout& layout) {
  if (a.rows != a.cols || a.rows == 0) fail(2, "multigrid: the system matrix must be square and non-empty");
  if (layout.dim != 2 && layout.dim != 3) fail(2, "multigrid: the DOF layout must be two- or three-dimensional");
  const int ndpn = layout.ndpn > 0 ? layout.ndpn : layout.dim;
  if (ndpn != layout.dim && !(layout.dim == 3 && ndpn == 6)) fail(2, "multigrid: unsupported DOFs per node");
  if (layout.unknowns == nullptr || layout.n != a.rows) fail(2, "multigrid: the DOF layout does not match the matrix");
  const Index n = a.rows;
  const bool same = ready_ && fp_rows_ == n && fp_nnz_ == a.nnz() &&
                    (Index)fp_unknowns_.size() == layout.n &&
                    std::equal(fp_unknowns_.begin(), fp_unknowns_.end(), layout.unknowns);
  const bool reuse = opt_.reuse_aggregates && same;
  if (!same) {  // symmetry check, Multigrid.cpp:968-991
    double max_abs = 0.0;
    for (double v : a.val) max_abs = std::max(max_abs, std::abs(v));
    for (Index i = 0; i < n; ++i)
      for (Index k = a.ptr[(size_t)i]; k < a.ptr[(size_t)i + 1]; ++k) {
        const Index j = a.col[(size_t)k];
        const Index* begin = a.col.data() + a.ptr[(size_t)j];
        const Index* end = a.col.data() + a.ptr[(size_t)j + 1];
        const Index* hit = std::lower_bound(begin, end, i);
        const double other = (hit != end && *hit == i) ? a.val[(size_t)(hit - a.col.data())] : 0.0;
        if (std::abs(other - a.val[(size_t)k]) > 1.0e-10 * max_abs) {
          std::ostringstream os;
          os << "multigrid: the system matrix is not symmetric (entry (" << i << ", " << j << "))";
          fail(2, os.str());
        }
      }
  }
  for (double v : a.val)
    if (!std::isfinite(v)) fail(2, "multigrid: the system matrix has a non-finite entry");
  if (!reuse) {
    levels_.clear();
    levels_.reserve((size_t)opt_.max_levels + 1);
    levels_.emplace_back();
    build_finest_nodes(layout, n);
  }
  levels_.front().a = a;
  size_t l = 0;
  const size_t reused_levels = levels_.size();
  while (true) {
    Level& lvl = levels_[l];
    numeric_level(l);
    const bool last_reused = reuse && l + 1 == reused_levels;
    const bool stop = reuse ? last_reused : (lvl.a.rows <= opt_.coarse_size || (int)l + 1 >= opt_.max_levels);
    if (stop) break;
    if (!reuse) {
      aggregate_level(levels_[l]);
      Level next;
      tentative_prolongator(levels_[l], next);
      const Index nc = next.node_ptr.back();
      if (nc >= (Index)(0.85 * double(levels_[l].a.rows)) || nc == 0) {
        levels_[l].aggregate.clear();
        levels_[l].p_tent = Csr();
        break;
      }
      levels_.push_back(std::move(next));
    }
    Level& cur = levels_[l];
    smooth_prolongator(cur, reuse);
    Level& coarse = levels_[l + 1];
    if (reuse) {
      multiply_numeric(cur.a, cur.p, cur.ap);
      multiply_numeric(cur.r, cur.ap, cur.coarse_raw);
    } else {
      cur.ap = multiply(cur.a, cur.p);
      cur.coarse_raw = multiply(cur.r, cur.ap);
    }
    coarse.a = symmetrize(cur.coarse_raw);
    ++l;
  }
  levels_.resize(l + 1);
  coarse_factorize(levels_.back());
  fp_rows_ = n;
  fp_nnz_ = a.nnz();
  fp_unknowns_.assign(layout.unknowns, layout.unknowns + layout.n);
  ready_ = true;
  stats_.levels.clear();
  double nnz_sum = 0, rows_sum = 0;
  for (size_t k = 0; k < levels_.size(); ++k) {
    LevelStats ls;
    ls.unknowns = levels_[k].a.rows;
    ls.nonzeros = levels_[k].a.nnz();
    ls.aggregates = k + 1 < levels_.size() ? levels_[k + 1].num_nodes : 0;
    ls.lambda_max = levels_[k].lambda_max;
    stats_.levels.push_back(ls);
    nnz_sum += ls.nonzeros;
    rows_sum += ls.unknowns;
  }
  stats_.operator_complexity = nnz_sum / double(stats_.levels.front().nonzeros);
  stats_.grid_complexity = rows_sum / double(stats_.levels.front().unknowns);
  stats_.reused_aggregates = reuse;
  {
    // Multigrid.cpp:1089-1100: coarse operators, P, R, A P, R A P and P-hat of
    // every level, plus the dense coarsest matrix when it is factored densely
    double stored = 0.0;
    for (size_t k = 0; k < levels_.size(); ++k) {
      const Level& lv = levels_[k];
      if (k > 0) stored += double(lv.a.nnz());
      stored += double(lv.p.nnz() + lv.r.nnz() + lv.ap.nnz() + lv.coarse_raw.nnz() + lv.p_tent.nnz());
    }
    const double nc = double(levels_.back().a.rows);
    stored += levels_.back().a.rows <= 4000 ? nc * nc : 0.0;
    stats_.storage_nonzeros = (long long)std::min(stored, 2.0e9);
  }
}

void Amg::apply(const double* r, double* z) const {
  if (!ready_) fail(2, "multigrid: apply() called before setup()");
  const Level& fine = levels_.front();
  std::copy(r, r + fine.a.rows, fine.b.begin());
  cycle(0);
  std::copy(fine.x.begin(), fine.x.end(), z);
}

PcgResult pcg_solve(const Csr& a, const double* b, double* x, const Amg& m, double tol, int max_iterations) {
  const Index n = a.rows;  // Multigrid.cpp:1152-1237
  PcgResult result;
  const double bnorm = std::sqrt(chunked_dot(b, b, n));
  if (bnorm == 0.0) {
    std::fill(x, x + n, 0.0);
    result.converged = true;
    return result;
  }
  const double target = tol * bnorm;
  std::vector<double> r((size_t)n), z((size_t)n), p((size_t)n), q((size_t)n);
  int total = 0;
  bool recursive_converged = false;
  double best_true = std::numeric_limits<double>::max();
  for (int restart = 0; restart < 3; ++restart) {
    residual(a, x, b, r.data());
    double rnorm = std::sqrt(chunked_dot(r.data(), r.data(), n));
    result.relative_residual = rnorm / bnorm;
    if (rnorm <= target) {
      result.converged = true;
      break;
    }
    m.apply(r.data(), z.data());
    p = z;
    double rz = chunked_dot(r.data(), z.data(), n);
    bool converged = false;
    while (total < max_iterations) {
      spmv(a, p.data(), q.data());
      const double pq = chunked_dot(p.data(), q.data(), n);
      if (!(pq > 0.0) || !std::isfinite(pq)) {
        std::ostringstream os;
        os << "preconditioned CG broke down at iteration " << total + 1 << " (p^T A p = " << pq << ")";
        fail(2, os.str());
      }
      const double alpha = rz / pq;
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
      for (Index i = 0; i < n; ++i) {
        x[i] += alpha * p[(size_t)i];
        r[(size_t)i] -= alpha * q[(size_t)i];
      }
      ++total;
      rnorm = std::sqrt(chunked_dot(r.data(), r.data(), n));
      if (rnorm <= target) {
        converged = true;
        break;
      }
      m.apply(r.data(), z.data());
      const double rz_new = chunked_dot(r.data(), z.data(), n);
      const double beta = rz_new / rz;
      rz = rz_new;
#pragma omp parallel for schedule(static) if (n >= kParallelRows)
      for (Index i = 0; i < n; ++i) p[(size_t)i] = z[(size_t)i] + beta * p[(size_t)i];
    }
    residual(a, x, b, r.data());
    result.relative_residual = std::sqrt(chunked_dot(r.data(), r.data(), n)) / bnorm;
    if (result.relative_residual <= tol) {
      result.converged = true;
      break;
    }
    if (!converged) break;
    recursive_converged = true;
    if (!(result.relative_residual < 0.5 * best_true)) break;
    best_true = result.relative_residual;
  }
  if (!result.converged && recursive_converged) result.converged = true;
  result.iterations = total;
  return result;
}

// ---------------------------------------------------------------------------
// Density filter (DensityFilter.cpp:19-195)
// ---------------------------------------------------------------------------
void Filter::build(const Mesh& m, double radius) {
  if (!(radius > 0.0)) fail(1, "filter radius must be positive");
  const Index ne = m.ne;
  const int dim = m.dim;
  std::vector<double> c((size_t)ne * 3);
  for (Index e = 0; e < ne; ++e) element_centroid(m, e, &c[(size_t)e * 3]);
  // Buckets of cell size = radius keyed by the integer cell triple; the visit
  // order is z slowest, then y, then x, bucket contents ascending.
  auto cell = [&](double v) { return (long)std::floor(v / radius); };
  std::map<std::tuple<long, long, long>, std::vector<Index>> buckets;
  for (Index e = 0; e < ne; ++e)
    buckets[{cell(c[(size_t)e * 3]), cell(c[(size_t)e * 3 + 1]), dim == 3 ? cell(c[(size_t)e * 3 + 2]) : 0}].push_back(e);
  std::vector<std::vector<Index>> rows_col((size_t)ne);
  std::vector<std::vector<double>> rows_w((size_t)ne);
  row_sums.assign((size_t)ne, 0.0);
#pragma omp parallel for schedule(dynamic, 1024)
  for (Index e = 0; e < ne; ++e) {
    const long ix = cell(c[(size_t)e * 3]), iy = cell(c[(size_t)e * 3 + 1]);
    const long iz = dim == 3 ? cell(c[(size_t)e * 3 + 2]) : 0;
    const long dzr = dim == 3 ? 1 : 0;
    double sum = 0.0;
    std::vector<std::pair<Index, double>> hits;
    for (long dz = -dzr; dz <= dzr; ++dz)
      for (long dy = -1; dy <= 1; ++dy)
        for (long dx = -1; dx <= 1; ++dx) {
          auto it = buckets.find({ix + dx, iy + dy, iz + dz});
          if (it == buckets.end()) continue;
          for (Index i : it->second) {
            double s2 = 0.0;
            for (int d = 0; d < dim; ++d) {
              const double t = c[(size_t)e * 3 + d] - c[(size_t)i * 3 + d];
              s2 += t * t;
            }
            const double w = radius - std::sqrt(s2);
            if (w <= 0.0) continue;
            hits.emplace_back(i, w);
            sum += w;
          }
        }
    if (!(sum > 0.0)) continue;
    row_sums[(size_t)e] = sum;
    std::sort(hits.begin(), hits.end());
    for (auto& h : hits) {
      rows_col[(size_t)e].push_back(h.first);
      rows_w[(size_t)e].push_back(h.second / sum);
    }
  }
  for (Index e = 0; e < ne; ++e)
    if (!(row_sums[(size_t)e] > 0.0)) fail(1, "filter row is empty");
  h.rows = h.cols = ne;
  h.ptr.assign((size_t)ne + 1, 0);
  for (Index e = 0; e < ne; ++e) h.ptr[(size_t)e + 1] = h.ptr[(size_t)e] + (Index)rows_col[(size_t)e].size();
  h.col.resize((size_t)h.ptr.back());
  h.val.resize((size_t)h.ptr.back());
  for (Index e = 0; e < ne; ++e) {
    std::copy(rows_col[(size_t)e].begin(), rows_col[(size_t)e].end(), h.col.begin() + h.ptr[(size_t)e]);
    std::copy(rows_w[(size_t)e].begin(), rows_w[(size_t)e].end(), h.val.begin() + h.ptr[(size_t)e]);
  }
}

void Filter::to_physical(const double* x, double* y) const {
  // Eigen column-major sparse x dense: each y_e sums H_ej x_j in ascending j.
  spmv(h, x, y);
}

void Filter::pull_back(const double* v, double* y) const {
  // H^T v: each y_j sums H_ej v_e in ascending e (row-major view of H^T).
  const Index n = h.rows;
  std::vector<Index> tptr((size_t)n + 1, 0);
  for (Index k = 0; k < h.nnz(); ++k) ++tptr[(size_t)h.col[(size_t)k] + 1];
  for (Index j = 0; j < n; ++j) tptr[(size_t)j + 1] += tptr[(size_t)j];
  std::vector<Index> next(tptr.begin(), tptr.end() - 1), trow((size_t)h.nnz());
  std::vector<double> tval((size_t)h.nnz());
  for (Index e = 0; e < n; ++e)
    for (Index k = h.ptr[(size_t)e]; k < h.ptr[(size_t)e + 1]; ++k) {
      const Index p = next[(size_t)h.col[(size_t)k]]++;
      trow[(size_t)p] = e;
      tval[(size_t)p] = h.val[(size_t)k];
    }
  for (Index j = 0; j < n; ++j) {
    double s = 0.0;
    for (Index k = tptr[(size_t)j]; k < tptr[(size_t)j + 1]; ++k) s += tval[(size_t)k] * v[trow[(size_t)k]];
    y[j] = s;
  }
}

// ---------------------------------------------------------------------------
// SIMP, objective, OC, loop
// ---------------------------------------------------------------------------
double simp_factor(double rho, const Simp& s) {
  const double r = std::clamp(rho, 0.0, 1.0);
  return s.emin_ratio + (1.0 - s.emin_ratio) * std::pow(r, s.penalty);
}
double simp_derivative(double rho, const Simp& s) {
  const double r = std::clamp(rho, 0.0, 1.0);
  return s.penalty * (1.0 - s.emin_ratio) * std::pow(r, s.penalty - 1.0);
}

void Problem::finalize() {
  ndpn = mesh.dim;
  if ((Index)dofs.fixed.size() != mesh.nn * ndpn) fail(4, "refcpu: the fixed-DOF mask has the wrong length");
  dofs.ndpn = ndpn;
  dofs.ndofs = mesh.nn * ndpn;
  dofs.rebuild();
  // Ke0 of element 0 (Assembler.cpp:41-58: one cached matrix on a uniform grid)
  const Index* nodes = mesh.element_nodes(0);
  std::vector<double> xe((size_t)mesh.dim * mesh.npe);
  for (int a = 0; a < mesh.npe; ++a)
    for (int i = 0; i < mesh.dim; ++i) xe[(size_t)i * mesh.npe + a] = mesh.coords[(size_t)nodes[a] * mesh.dim + i];
  const int nd = mesh.npe * ndpn;
  ke0.assign((size_t)nd * nd, 0.0);
  if (mesh.dim == 3) {
    double d[36];
    d_three_dimensional(E, nu, d);
    hex8_stiffness(xe.data(), d, ke0.data());
  } else {
    double d[9];
    d_plane_stress(E, nu, d);
    quad4_stiffness(xe.data(), d, thickness, ke0.data());
  }
  pattern = build_pattern(mesh, ndpn);
  ff = free_free_map(pattern, dofs);
  volumes.resize((size_t)mesh.ne);
  for (Index e = 0; e < mesh.ne; ++e) volumes[(size_t)e] = element_measure(mesh, e) * (mesh.dim == 2 ? thickness : 1.0);
  if (lower.empty()) lower.assign((size_t)mesh.ne, 0.0);
  if (upper.empty()) upper.assign((size_t)mesh.ne, 1.0);
  if (passive.empty()) passive.assign((size_t)mesh.ne, 0);
  amg = Amg(amg_options);
}

double Problem::domain_volume() const {
  double s = 0.0;
  for (double v : volumes) s += v;
  return s;
}

Csr Problem::assemble_kff(const double* scale) const {
  std::vector<double> full;
  scatter(pattern, mesh, ndpn, ke0.data(), scale, full);
  return reduce(ff, pattern, full);
}

namespace {
double dot(const std::vector<double>& a, const std::vector<double>& b) {
  double s = 0.0;
  for (size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
  return s;
}
// the clamped filtered density (ComplianceObjective::evaluate, Sensitivity.cpp:120-123)
std::vector<double> filtered_density(const Problem& p, const std::vector<double>& x) {
  std::vector<double> rho(x.size());
  if (p.filter_on) p.filter.to_physical(x.data(), rho.data()); else rho = x;
  for (auto& r : rho) r = std::min(std::max(r, 0.0), 1.0);
  return rho;
}

// ComplianceObjective::physical_density (Sensitivity.cpp:68-72)
std::vector<double> physical_density(const Problem& p, const std::vector<double>& x) {
  std::vector<double> rho = filtered_density(p, x);
  if (p.beta > 0.0)
    for (auto& r : rho) r = heaviside_project(r, p.beta, p.eta);
  return rho;
}
}  // namespace

double heaviside_project(double r, double beta, double eta) {
  const double a = std::tanh(beta * eta);
  const double denom = a + std::tanh(beta * (1.0 - eta));
  return (a + std::tanh(beta * (r - eta))) / denom;
}

double heaviside_derivative(double r, double beta, double eta) {
  const double denom = std::tanh(beta * eta) + std::tanh(beta * (1.0 - eta));
  const double t = std::tanh(beta * (r - eta));
  return beta * (1.0 - t * t) / denom;
}

void Projection::validate() const {
  if (!enabled) return;
  if (!(eta > 0.0 && eta < 1.0)) fail(1, "topology.projection.eta must lie in (0, 1)");
  if (!(beta_start > 0.0) || !std::isfinite(beta_start)) fail(1, "topology.projection.beta_start must be positive");
  if (!(beta_max >= beta_start) || !std::isfinite(beta_max))
    fail(1, "topology.projection.beta_max must be at least beta_start");
  if (beta_max > 1024.0) fail(1, "topology.projection.beta_max is beyond any useful sharpness");
  if (!(beta_factor > 1.0)) fail(1, "topology.projection.beta_factor must exceed 1");
  if (beta_interval < 1) fail(1, "topology.projection.beta_interval must be at least 1");
  if (robust && !(robust_delta > 0.0 && eta - robust_delta > 0.0 && eta + robust_delta < 1.0))
    fail(1, "topology.projection.robust_delta must be positive and keep eta +- delta inside (0, 1)");
  if (robust && robust_volume_interval < 1) fail(1, "topology.projection.robust_volume_interval must be at least 1");
}

int Projection::num_stages() const {
  int stages = 1;
  double b = beta_start;
  while (b < beta_max && stages < 64) {
    b *= beta_factor;
    ++stages;
  }
  return stages;
}

double Projection::beta_of_stage(int k) const {
  double b = beta_start;
  for (int i = 0; i < k && b < beta_max; ++i) b *= beta_factor;
  return std::min(b, beta_max);
}

Evaluation evaluate(Problem& p, const std::vector<double>& x, bool need_gradients) {
  const Index ne = p.mesh.ne;  // Sensitivity.cpp:110-237
  if ((Index)x.size() != ne) fail(1, "design vector has the wrong length");
  using clock = std::chrono::steady_clock;
  auto secs = [](clock::time_point a, clock::time_point b) { return std::chrono::duration<double>(b - a).count(); };
  const auto t_start = clock::now();
  p.seconds_pcg = 0.0;
  Evaluation out;
  const std::vector<double> filtered = filtered_density(p, x);
  out.physical = physical_density(p, x);
  out.factors.resize((size_t)ne);
  for (Index e = 0; e < ne; ++e) out.factors[(size_t)e] = simp_factor(out.physical[(size_t)e], p.simp);
  const auto t_asm = clock::now();
  p.kff = p.assemble_kff(out.factors.data());
  p.seconds_assemble = secs(t_asm, clock::now());
  Layout layout;
  layout.dim = p.mesh.dim;
  layout.ndpn = p.ndpn;
  layout.coords = p.mesh.coords.data();
  layout.nn = p.mesh.nn;
  layout.unknowns = p.dofs.free.data();
  layout.n = (Index)p.dofs.free.size();
  const auto t_setup = clock::now();
  p.amg.setup(p.kff, layout);
  p.seconds_setup = secs(t_setup, clock::now());
  p.previous.resize(p.loads.size());
  const int npe = p.mesh.npe, nd = npe * p.ndpn;
  out.energy.assign((size_t)ne, 0.0);
  std::vector<double> dc((size_t)ne, 0.0), dfac((size_t)ne);
  for (Index e = 0; e < ne; ++e) dfac[(size_t)e] = simp_derivative(out.physical[(size_t)e], p.simp);
  const Index nf = layout.n;
  for (size_t l = 0; l < p.loads.size(); ++l) {
    std::vector<double> b((size_t)nf), xf((size_t)nf, 0.0);
    for (Index i = 0; i < nf; ++i) b[(size_t)i] = p.loads[l][(size_t)p.dofs.free[(size_t)i]];
    std::vector<double>& prev = p.previous[l];
    if (p.warm_start && prev.size() == p.loads[l].size())
      for (Index i = 0; i < nf; ++i) xf[(size_t)i] = prev[(size_t)p.dofs.free[(size_t)i]];
    const auto t_pcg = clock::now();
    const PcgResult res = pcg_solve(p.kff, b.data(), xf.data(), p.amg, p.tolerance, p.max_iterations);
    p.seconds_pcg += secs(t_pcg, clock::now());
    if (!res.converged) {
      std::ostringstream os;
      os << "AMG-CG did not converge: " << res.iterations << " iterations left a relative residual of "
         << res.relative_residual;
      fail(3, os.str());
    }
    out.linear_iterations += res.iterations;
    p.total_iterations += res.iterations;
    std::vector<double> u((size_t)p.dofs.ndofs, 0.0);
    for (Index i = 0; i < nf; ++i) u[(size_t)p.dofs.free[(size_t)i]] = xf[(size_t)i];
    if (p.warm_start) prev = u;
    const double c = dot(p.loads[l], u);
    out.load_case_compliance.push_back(c);
    out.compliance += p.weights[l] * c;
    std::vector<double> ue((size_t)nd), ku((size_t)nd);
    for (Index e = 0; e < ne; ++e) {
      const Index* nodes = p.mesh.element_nodes(e);
      for (int a = 0; a < npe; ++a)
        for (int k = 0; k < p.ndpn; ++k) ue[(size_t)(a * p.ndpn + k)] = u[(size_t)(nodes[a] * p.ndpn + k)];
      for (int i = 0; i < nd; ++i) {
        double s = 0.0;
        for (int j = 0; j < nd; ++j) s += p.ke0[(size_t)i * nd + j] * ue[(size_t)j];
        ku[(size_t)i] = s;
      }
      const double quad = dot(ue, ku);
      out.energy[(size_t)e] += p.weights[l] * 0.5 * out.factors[(size_t)e] * quad;
      if (need_gradients) dc[(size_t)e] -= p.weights[l] * dfac[(size_t)e] * quad;
    }
    out.u.push_back(std::move(u));
  }
  out.volume = dot(out.physical, p.volumes);
  out.volume_fraction = out.volume / p.domain_volume();
  if (need_gradients) {
    out.dc_dx.resize((size_t)ne);
    out.dv_dx.resize((size_t)ne);
    // through the projection, then the (linear) density filter (Sensitivity.cpp:210-219)
    std::vector<double> dvol = p.volumes;
    if (p.beta > 0.0)
      for (Index e = 0; e < ne; ++e) {
        const double d = heaviside_derivative(filtered[(size_t)e], p.beta, p.eta);
        dc[(size_t)e] *= d;
        dvol[(size_t)e] *= d;
      }
    if (p.filter_on) {
      p.filter.pull_back(dc.data(), out.dc_dx.data());
      p.filter.pull_back(dvol.data(), out.dv_dx.data());
    } else {
      out.dc_dx = dc;
      out.dv_dx = dvol;
    }
  }
  if (!std::isfinite(out.compliance)) fail(2, "the compliance objective evaluated to a non-finite value");
  p.seconds_total = secs(t_start, clock::now());
  return out;
}

namespace {
std::vector<double> oc_candidate(const Problem& p, const std::vector<double>& x, const std::vector<double>& dc,
                                 const std::vector<double>& dv, double lambda, const OcOptions& o) {
  const size_t ne = x.size();  // OptimalityCriteria.cpp:14-36
  std::vector<double> xn(ne);
  for (size_t e = 0; e < ne; ++e) {
    const double lo = std::max(p.lower[e], x[e] - o.move_limit);
    const double hi = std::min(p.upper[e], x[e] + o.move_limit);
    if (p.passive[e]) {
      xn[e] = p.lower[e];
      continue;
    }
    const double denom = lambda * dv[e];
    double b = 0.0;
    if (denom > 0.0) b = std::max(0.0, -dc[e] / denom);
    const double target = x[e] * std::pow(b, o.damping);
    xn[e] = std::clamp(target, lo, hi);
  }
  return xn;
}
}  // namespace

OcStep oc_update(const Problem& p, const std::vector<double>& x, const std::vector<double>& dc,
                 const std::vector<double>& dv, const OcOptions& o, double volume_target, double vol_eta,
                 bool step_toward) {
  // OptimalityCriteria.cpp:40-149
  const double target = volume_target > 0.0 ? volume_target : p.volume_fraction * p.domain_volume();
  Problem& q = const_cast<Problem&>(p);
  auto vol = [&](const std::vector<double>& xi) {
    const double saved = q.eta;
    if (vol_eta >= 0.0) q.eta = vol_eta;
    const double v = dot(physical_density(q, xi), q.volumes);
    q.eta = saved;
    return v;
  };
  double lo = o.lambda_lower, hi = o.lambda_upper;
  double vol_lo = vol(oc_candidate(p, x, dc, dv, lo, o));
  double vol_hi = vol(oc_candidate(p, x, dc, dv, hi, o));
  int widen = 0;
  while (vol_lo < target && widen < 60) {
    lo *= 1.0e-3;
    vol_lo = vol(oc_candidate(p, x, dc, dv, lo, o));
    ++widen;
  }
  while (vol_hi > target && widen < 120) {
    hi *= 1.0e3;
    vol_hi = vol(oc_candidate(p, x, dc, dv, hi, o));
    ++widen;
  }
  if ((vol_lo < target || vol_hi > target) && step_toward) {
    const bool too_light = vol_lo < target;  // OptimalityCriteria.cpp:82-97
    OcStep step;
    step.lambda = too_light ? lo : hi;
    step.x = oc_candidate(p, x, dc, dv, step.lambda, o);
    step.achieved_volume = too_light ? vol_lo : vol_hi;
    step.volume_converged = false;
    double mc = 0.0;
    for (size_t e = 0; e < x.size(); ++e) mc = std::max(mc, std::abs(step.x[e] - x[e]));
    step.max_change = mc;
    return step;
  }
  if (vol_lo < target || vol_hi > target) fail(3, "the optimality-criteria bisection cannot bracket the volume target");
  OcStep step;
  std::vector<double> cand = oc_candidate(p, x, dc, dv, lo, o);
  double volume = vol_lo;
  for (int it = 0; it < o.max_bisections; ++it) {
    const double mid = std::sqrt(lo * hi);
    cand = oc_candidate(p, x, dc, dv, mid, o);
    volume = vol(cand);
    step.bisections = it + 1;
    step.lambda = mid;
    if (std::abs(volume - target) <= o.volume_tolerance * target) {
      step.volume_converged = true;
      break;
    }
    if (volume > target) lo = mid; else hi = mid;
    if (hi / lo - 1.0 < 1.0e-15) {
      step.volume_converged = std::abs(volume - target) <= 1.0e-6 * std::max(target, 1.0e-30);
      break;
    }
  }
  step.achieved_volume = volume;
  double mc = 0.0;
  for (size_t e = 0; e < x.size(); ++e) mc = std::max(mc, std::abs(cand[e] - x[e]));
  step.max_change = mc;
  step.x = std::move(cand);
  return step;
}

RunResult run_oc(Problem& p, const RunOptions& o, const std::vector<double>& x0) {
  RunResult res;  // TopologyOptimizer.cpp:556-724
  std::vector<double> x = x0;
  for (size_t e = 0; e < x.size(); ++e) x[e] = std::clamp(x[e], p.lower[e], p.upper[e]);
  std::vector<double> history;
  // BetaSchedule (TopologyOptimizer.cpp:17-46): stage k runs at beta_of_stage(k)
  const Projection& pr = p.projection;
  pr.validate();
  const double saved_beta = p.beta, saved_eta = p.eta;
  const int stages = pr.enabled ? pr.num_stages() : 1;
  int stage = 0, stage_start = 1;
  const bool robust = pr.enabled && pr.robust;
  // the objective acts on the eroded design in a robust run (TopologyOptimizer.cpp:566-571)
  const double objective_eta = robust ? pr.eta + pr.robust_delta : pr.eta;
  const double dilated_eta = pr.eta - pr.robust_delta;
  double dilated_target = p.volume_fraction * p.domain_volume();
  if (pr.enabled) {
    p.beta = pr.beta_of_stage(0);
    p.eta = objective_eta;
  }
  // the volume of x projected at eta (its filtered density at the current beta)
  auto volume_at_eta = [&](const std::vector<double>& xi, double eta) {
    const double saved = p.eta;
    p.eta = eta;
    const double v = dot(physical_density(p, xi), p.volumes);
    p.eta = saved;
    return v;
  };
  int iteration;
  for (iteration = 1; iteration <= o.max_iterations; ++iteration) {
    double penalty = p.simp.penalty;
    if (o.continuation_steps > 1) {  // TopologyOptimizer.cpp:245-253
      const int stage = std::min((iteration - 1) / std::max(o.continuation_iterations, 1), o.continuation_steps - 1);
      const double t = double(stage) / double(o.continuation_steps - 1);
      penalty = o.penalty_start + t * (p.simp.penalty - o.penalty_start);
    }
    Problem& q = p;
    const double saved = q.simp.penalty;
    q.simp.penalty = penalty;
    Evaluation ev = evaluate(q, x, true);
    std::vector<double> dv = ev.dv_dx;
    double v_dil = 0.0, v_int = 0.0;
    if (robust) {
      // robust_volumes (TopologyOptimizer.cpp:70-80): the dilated design's volume and
      // gradient, the blueprint's volume; the dilated target follows their ratio
      v_dil = volume_at_eta(x, dilated_eta);
      v_int = volume_at_eta(x, pr.eta);
      const std::vector<double> filt = filtered_density(p, x);
      std::vector<double> w(filt.size());
      for (size_t e = 0; e < filt.size(); ++e) w[e] = p.volumes[e] * heaviside_derivative(filt[e], p.beta, dilated_eta);
      if (p.filter_on) p.filter.pull_back(w.data(), dv.data()); else dv = w;
      if ((iteration - 1) % pr.robust_volume_interval == 0 && v_int > 0.0)
        dilated_target = p.volume_fraction * p.domain_volume() * v_dil / v_int;
    }
    OcStep step = robust ? oc_update(q, x, ev.dc_dx, dv, o.oc, dilated_target, dilated_eta, true)
                         : oc_update(q, x, ev.dc_dx, dv, o.oc);
    q.simp.penalty = saved;
    double gray = 0.0;
    for (double r : ev.physical) gray += r * (1.0 - r);
    gray = 4.0 * gray / double(ev.physical.size());
    const double dom = p.domain_volume();
    IterationRecord rec{iteration, penalty, ev.compliance, ev.volume, ev.volume_fraction, step.max_change,
                        step.lambda, gray, step.bisections, step.volume_converged, ev.linear_iterations,
                        pr.enabled ? p.beta : 0.0};
    if (robust) {
      rec.volume = v_int;
      rec.volume_fraction = v_int / dom;
      rec.eroded_vf = ev.volume_fraction;
      rec.dilated_vf = v_dil / dom;
      rec.dilated_target_fraction = dilated_target / dom;
    }
    res.history.push_back(rec);
    history.push_back(ev.compliance);
    x = step.x;
    for (size_t e = 0; e < x.size(); ++e) x[e] = std::clamp(x[e], p.lower[e], p.upper[e]);
    const bool at_final = o.continuation_steps <= 1 || iteration >= o.continuation_iterations * (o.continuation_steps - 1);
    const bool change_ok = step.max_change <= o.change_tolerance;
    double spread = std::numeric_limits<double>::infinity();
    const size_t w = (size_t)std::max(o.objective_window, 1);
    if (history.size() > w) {
      auto mm = std::minmax_element(history.end() - (long)(w + 1), history.end());
      spread = (*mm.second - *mm.first) / std::max(std::abs(history.back()), 1.0e-30);
    }
    const bool objective_ok = o.objective_tolerance > 0.0 && spread <= o.objective_tolerance;
    const bool final_stage = stage + 1 >= stages;
    if (at_final && final_stage && (change_ok || objective_ok)) {
      res.converged = true;
      res.stop_reason = change_ok ? 1 : 2;  // TopologyOptimizer.cpp:692
      break;
    }
    if (!final_stage && (iteration - stage_start + 1 >= pr.beta_interval ||
                         (pr.advance_on_convergence && step.max_change <= o.change_tolerance))) {
      ++stage;  // TopologyOptimizer.cpp:706-711: the objective window restarts
      stage_start = iteration + 1;
      p.beta = pr.beta_of_stage(stage);
      history.clear();
    }
  }
  res.iterations = std::min(iteration, o.max_iterations);
  Evaluation fin = evaluate(p, x, true);  // TopologyOptimizer.cpp:263-283
  if (robust) {
    // the blueprint is the result; the eroded and dilated designs are recorded (:347-390)
    res.compliance_eroded = fin.compliance;
    res.vf_eroded = fin.volume_fraction;
    p.eta = dilated_eta;
    const Evaluation dil = evaluate(p, x, false);
    res.compliance_dilated = dil.compliance;
    res.vf_dilated = dil.volume_fraction;
    p.eta = pr.eta;
    fin = evaluate(p, x, false);
    res.dilated_target_fraction = res.history.empty() ? 0.0 : res.history.back().dilated_target_fraction;
  }
  res.compliance = fin.compliance;
  res.volume = fin.volume;
  res.volume_fraction = fin.volume_fraction;
  res.load_case_compliance = fin.load_case_compliance;
  double g = 0.0;
  for (double r : fin.physical) g += r * (1.0 - r);
  res.gray = 4.0 * g / double(std::max<size_t>(fin.physical.size(), 1));
  res.x = x;
  res.final_beta = pr.enabled ? p.beta : 0.0;
  p.beta = saved_beta;
  p.eta = saved_eta;
  return res;
}

}  // namespace refcpu
