// serac_oracle_form.hpp -- CPU FP64 restatement of serac::Functional's element
// loop (residual, linearize, J.v, element gradients) and of the Newton /
// Jacobi-CG loop.  TEST INFRASTRUCTURE ONLY (see serac_oracle.hpp).
#pragma once

#ifdef _OPENMP
#include <omp.h>
#endif
#include <array>
#include <map>

#include "serac_oracle.hpp"

namespace oracle {

// H1 parent <-> physical maps, fn/finite_element.hpp:246-322.
//   J[row][col] = d x_row / d xi_col  (= jacobians(col,row,q), :259-264)
//   grad_x u = grad_xi u . inv(J);  source *= detJ;  flux = flux . inv(J^T) detJ
template <typename T>
inline void parent_to_physical(int nc, const double Jm[3][3], T* grad)
{
  M3<double> J;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) J.a[r][c] = Jm[r][c];
  M3<double> Ji = inv3(J);
  for (int c = 0; c < nc; c++) {
    T g[3] = {grad[c * 3 + 0], grad[c * 3 + 1], grad[c * 3 + 2]};
    for (int k = 0; k < 3; k++) grad[c * 3 + k] = g[0] * Ji.a[0][k] + g[1] * Ji.a[1][k] + g[2] * Ji.a[2][k];
  }
}

template <typename T>
inline void physical_to_parent(int nc, const double Jm[3][3], T* source, T* flux)
{
  M3<double> JT;
  for (int r = 0; r < 3; r++)
    for (int c = 0; c < 3; c++) JT.a[r][c] = Jm[c][r];
  double     dv  = det3(JT);
  M3<double> JTi = inv3(JT);
  for (int c = 0; c < nc; c++) {
    source[c] = source[c] * dv;
    T f[3]    = {flux[c * 3 + 0], flux[c * 3 + 1], flux[c * 3 + 2]};
    for (int j = 0; j < 3; j++)
      flux[c * 3 + j] = (f[0] * JTi.a[0][j] + f[1] * JTi.a[1][j] + f[2] * JTi.a[2][j]) * dv;
  }
}

struct Integral {
  QFParams         qf;
  std::vector<int> elems;
  bool             blk[2][2];
  int              dsize = 0;  // doubles per quadrature point of stored derivative
  std::vector<double> D;       // [elem-in-domain][q][dsize]
  // QuadratureData<State> (fn/quadrature_data.hpp:111-167): ssize doubles per
  // quadrature point, [elem-in-domain][q][ssize]; empty when stateless
  int                 ssize = 0;
  std::vector<double> state;
  // DependsOn<0, 1>: the integrand also reads the value of a second trial
  // argument (qf_eval2), taken from Form::U1 during evaluation
  bool                two = false;
};

// AddBoundaryIntegral (fn/functional.hpp:328-354): faces (element, local face
// 2 * axis + side) that belong to one element.  D holds the forward-mode
// tangent d(source * scale)/d[u, du/dxi] per face quadrature point.
struct BIntegral {
  QFParams                        qf;
  std::vector<std::array<int, 2>> faces;
  std::vector<double>             D;  // [face][q][nc][3 nc]
};

// Functional<test(trial)> with test == trial space, one trial argument.
// Restates fn/functional.hpp:406-498 (operator(), ActionOfGradient) and
// fn/domain_integral_kernels.hpp:155-304 (evaluation_kernel_impl,
// action_of_gradient_kernel), :327-361 (element_gradient_kernel).
struct Form {
  const Mesh*           mesh = nullptr;
  Space                 space;
  int                   Q = 2;
  Basis1D               basis;
  std::vector<double>   Xq, Jq;
  std::vector<Integral> integrals;
  std::vector<BIntegral> bintegrals;
  bool                  update_qdata = false;  // updateQdata(bool), fn/functional.hpp:523
  const double*         U1 = nullptr;          // the second trial argument of two-argument integrals

  Form(const Mesh& m, const Space& s) : mesh(&m), space(s), Q(s.p + 1), basis(s.p + 1, s.p + 1)
  {
    geometric_factors(m, Q, Xq, Jq);
  }

  int nq() const { return Q * Q * Q; }
  int nc() const { return space.nc; }

  // state_init: qf_state_size(kind) doubles every quadrature point starts from
  void add_integral(const QFParams& qf, const std::vector<int>& elems, const double* state_init = nullptr)
  {
    Integral I;
    I.qf    = qf;
    I.elems = elems;
    if (I.elems.empty())
      for (int e = 0; e < mesh->ne; e++) I.elems.push_back(e);
    qf_blocks(qf.kind, I.blk);
    int c    = nc();
    I.dsize  = (I.blk[0][0] ? c * c : 0) + (I.blk[0][1] ? c * 3 * c : 0) + (I.blk[1][0] ? 3 * c * c : 0) +
              (I.blk[1][1] ? 9 * c * c : 0);
    I.ssize = qf_state_size(qf.kind);
    if (I.ssize) {
      I.state.assign(I.elems.size() * size_t(nq()) * I.ssize, 0.0);
      if (state_init)
        for (size_t k = 0; k < I.state.size(); k++) I.state[k] = state_init[k % size_t(I.ssize)];
    }
    integrals.push_back(std::move(I));
  }

  // boundary faces on the bounding-box planes chosen by face_mask (bit 2d:
  // axis d at its minimum, bit 2d+1: at its maximum; 0: the whole boundary)
  void add_boundary_integral(const QFParams& qf, int face_mask)
  {
    if (bqf_components(qf.kind) != nc()) throw std::runtime_error("boundary qfunction / space components");
    std::map<std::array<int, 4>, int> count;
    auto face_verts = [&](int e, int lf) {
      std::array<int, 4> v;
      int                k = 0;
      for (int c = 0; c < 8; c++)
        if (((c >> (lf / 2)) & 1) == lf % 2) v[size_t(k++)] = mesh->hex[size_t(e) * 8 + size_t(LEX2V[c])];
      std::sort(v.begin(), v.end());
      return v;
    };
    for (int e = 0; e < mesh->ne; e++)
      for (int lf = 0; lf < 6; lf++) count[face_verts(e, lf)]++;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int v = 0; v < mesh->nv; v++)
      for (int d = 0; d < 3; d++) {
        lo[d] = std::min(lo[d], mesh->X[size_t(3 * v + d)]);
        hi[d] = std::max(hi[d], mesh->X[size_t(3 * v + d)]);
      }
    BIntegral B;
    B.qf = qf;
    for (int e = 0; e < mesh->ne; e++)
      for (int lf = 0; lf < 6; lf++) {
        auto v = face_verts(e, lf);
        if (count[v] != 1) continue;
        bool take = face_mask == 0;
        for (int bit = 0; bit < 6 && !take; bit++) {
          if (!((face_mask >> bit) & 1)) continue;
          int    d     = bit / 2;
          double plane = (bit & 1) ? hi[d] : lo[d], tol = 1e-12 * std::max(hi[d] - lo[d], 1.0);
          bool   on    = true;
          for (int x : v) on = on && std::fabs(mesh->X[size_t(3 * x + d)] - plane) <= tol;
          take = on;
        }
        if (take) B.faces.push_back({e, lf});
      }
    bintegrals.push_back(std::move(B));
  }

  // Evaluates boundary face (e, lf) on element DOFs Ue -> Re (added); with D,
  // also the tangent, or (jvp) the tangent applied to Ue.
  //   mode 0: residual, 1: residual + tangent into D, 2: J.v with D
  // The face is parametrized by its two tangential parent axes in increasing
  // order, swapped when dX/dxi1 x dX/dxi2 does not point away from the
  // element's centroid.
  void face_eval(const QFParams& qf, int e, int lf, const double* Ue, double* Re, double* D, int mode) const
  {
    const int c = nc(), n = space.p + 1, Qn = Q;
    const int d = lf / 2, side = lf % 2;
    int       ax[2] = {d == 0 ? 1 : 0, d == 2 ? 1 : 2};
    double    Xc[8][3];
    for (int k = 0; k < 8; k++) {
      const int v = mesh->hex[size_t(e) * 8 + size_t(LEX2V[k])];
      for (int i = 0; i < 3; i++) Xc[k][i] = mesh->X[size_t(3 * v + i)];
    }
    // trilinear map and its parent derivatives at xi
    auto geo = [&](const double xi[3], double X[3], double dXdxi[3][3]) {
      for (int i = 0; i < 3; i++) {
        X[i] = 0;
        for (int j = 0; j < 3; j++) dXdxi[i][j] = 0;
      }
      for (int k = 0; k < 8; k++) {
        double N[3], dN[3];
        for (int j = 0; j < 3; j++) {
          const int b = (k >> j) & 1;
          N[j]        = b ? xi[j] : 1.0 - xi[j];
          dN[j]       = b ? 1.0 : -1.0;
        }
        for (int i = 0; i < 3; i++) {
          X[i] += Xc[k][i] * N[0] * N[1] * N[2];
          dXdxi[i][0] += Xc[k][i] * dN[0] * N[1] * N[2];
          dXdxi[i][1] += Xc[k][i] * N[0] * dN[1] * N[2];
          dXdxi[i][2] += Xc[k][i] * N[0] * N[1] * dN[2];
        }
      }
    };
    {
      double xi[3] = {0.5, 0.5, 0.5}, X0[3], Xf[3], J0[3][3], Jf[3][3];
      geo(xi, X0, J0);
      xi[d] = side;
      geo(xi, Xf, Jf);
      const double* a = nullptr;
      (void)a;
      double t1[3] = {Jf[0][ax[0]], Jf[1][ax[0]], Jf[2][ax[0]]}, t2[3] = {Jf[0][ax[1]], Jf[1][ax[1]], Jf[2][ax[1]]};
      double nrm[3] = {t1[1] * t2[2] - t1[2] * t2[1], t1[2] * t2[0] - t1[0] * t2[2], t1[0] * t2[1] - t1[1] * t2[0]};
      double out    = nrm[0] * (Xf[0] - X0[0]) + nrm[1] * (Xf[1] - X0[1]) + nrm[2] * (Xf[2] - X0[2]);
      if (out < 0) std::swap(ax[0], ax[1]);
    }
    const int M = 3 * c;
    for (int q2 = 0; q2 < Qn; q2++)
      for (int q1 = 0; q1 < Qn; q1++) {
        double xi[3];
        xi[d]     = side;
        xi[ax[0]] = basis.xq[size_t(q1)];
        xi[ax[1]] = basis.xq[size_t(q2)];
        double X[3], J[3][3];
        geo(xi, X, J);
        double dX[3][2];
        for (int i = 0; i < 3; i++) dX[i][0] = J[i][ax[0]], dX[i][1] = J[i][ax[1]];
        double cr[3] = {dX[1][0] * dX[2][1] - dX[2][0] * dX[1][1], dX[2][0] * dX[0][1] - dX[0][0] * dX[2][1],
                        dX[0][0] * dX[1][1] - dX[1][0] * dX[0][1]};
        const double scale = std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]);
        const double w     = basis.wq[size_t(q1)] * basis.wq[size_t(q2)];
        // element basis and its derivatives along the two face axes
        std::vector<double> Bx(static_cast<size_t>(3 * n)), Gx(static_cast<size_t>(3 * n));
        for (int j = 0; j < 3; j++) {
          gll_basis(n, xi[j], &Bx[size_t(j * n)]);
          gll_dbasis(n, xi[j], &Gx[size_t(j * n)]);
        }
        const int           npe = n * n * n;
        std::vector<double> Nv(static_cast<size_t>(npe)), N1(static_cast<size_t>(npe)), N2(static_cast<size_t>(npe));
        for (int iz = 0; iz < n; iz++)
          for (int iy = 0; iy < n; iy++)
            for (int ix = 0; ix < n; ix++) {
              const int    i = (iz * n + iy) * n + ix;
              const int    id[3] = {ix, iy, iz};
              double       v = 1, g1 = 1, g2 = 1;
              for (int j = 0; j < 3; j++) {
                const double b = Bx[size_t(j * n + id[j])], g = Gx[size_t(j * n + id[j])];
                v *= b;
                g1 *= (j == ax[0]) ? g : b;
                g2 *= (j == ax[1]) ? g : b;
              }
              Nv[size_t(i)] = v, N1[size_t(i)] = g1, N2[size_t(i)] = g2;
            }
        double in[9];
        for (int cc = 0; cc < c; cc++) {
          double v = 0, g1 = 0, g2 = 0;
          for (int i = 0; i < npe; i++) {
            const double U = Ue[size_t(cc) * npe + size_t(i)];
            v += Nv[size_t(i)] * U;
            g1 += N1[size_t(i)] * U;
            g2 += N2[size_t(i)] * U;
          }
          in[cc] = v, in[c + 2 * cc] = g1, in[c + 2 * cc + 1] = g2;
        }
        double     out[3];
        const int  q  = q2 * Qn + q1;
        double*    Dq = D ? D + size_t(q) * size_t(c * M) : nullptr;
        if (mode == 2) {
          for (int r = 0; r < c; r++) {
            out[r] = 0;
            for (int s2 = 0; s2 < M; s2++) out[r] += Dq[r * M + s2] * in[s2];
          }
        } else {
          double src[3];
          bqf_eval(qf, X, dX, in, in + c, src);
          for (int r = 0; r < c; r++) out[r] = src[r] * scale;
          if (mode == 1) {
            dual<9> din[9], ds[3];
            for (int s2 = 0; s2 < M; s2++) {
              din[s2]              = dual<9>(in[s2]);
              din[s2].gradient[s2] = 1.0;
            }
            bqf_eval(qf, X, dX, din, din + c, ds);
            for (int r = 0; r < c; r++)
              for (int s2 = 0; s2 < M; s2++) Dq[r * M + s2] = (ds[r] * scale).gradient[size_t(s2)];
          }
        }
        for (int cc = 0; cc < c; cc++)
          for (int i = 0; i < npe; i++) Re[size_t(cc) * npe + size_t(i)] += w * Nv[size_t(i)] * out[cc];
      }
  }

  void boundary_eval(const double* U, double* R, bool store_D)
  {
    const int           c = nc(), npe = space.npe, nq2 = Q * Q;
    std::vector<double> Ue(static_cast<size_t>(c * npe)), Re(static_cast<size_t>(c * npe));
    for (auto& B : bintegrals) {
      if (store_D) B.D.assign(B.faces.size() * size_t(nq2) * size_t(c * 3 * c), 0.0);
      for (size_t k = 0; k < B.faces.size(); k++) {
        const int e = B.faces[k][0];
        gather_elem(e, U, Ue.data());
        std::fill(Re.begin(), Re.end(), 0.0);
        face_eval(B.qf, e, B.faces[k][1], Ue.data(), Re.data(),
                  store_D ? &B.D[k * size_t(nq2) * size_t(c * 3 * c)] : nullptr, store_D ? 1 : 0);
        scatter_elem(e, Re.data(), R);
      }
    }
  }

  void boundary_jvp(const double* dU, double* dR) const
  {
    const int           c = nc(), npe = space.npe, nq2 = Q * Q;
    std::vector<double> Ue(static_cast<size_t>(c * npe)), Re(static_cast<size_t>(c * npe));
    for (auto& B : bintegrals) {
      if (B.D.empty()) throw std::runtime_error("boundary jvp before linearize");
      for (size_t k = 0; k < B.faces.size(); k++) {
        const int e = B.faces[k][0];
        gather_elem(e, dU, Ue.data());
        std::fill(Re.begin(), Re.end(), 0.0);
        face_eval(B.qf, e, B.faces[k][1], Ue.data(), Re.data(),
                  const_cast<double*>(&B.D[k * size_t(nq2) * size_t(c * 3 * c)]), 2);
        scatter_elem(e, Re.data(), dR);
      }
    }
  }

  // the boundary face's matrix over its element's DOFs [(c,i)][(c',j)]
  void face_matrix(int bintegral, size_t k, double* K) const
  {
    const BIntegral& B  = bintegrals[size_t(bintegral)];
    const int        c = nc(), npe = space.npe, nd = c * npe, nq2 = Q * Q;
    std::vector<double> Ue(static_cast<size_t>(nd)), Re(static_cast<size_t>(nd));
    for (int col = 0; col < nd; col++) {
      std::fill(Ue.begin(), Ue.end(), 0.0);
      Ue[size_t(col)] = 1.0;
      std::fill(Re.begin(), Re.end(), 0.0);
      face_eval(B.qf, B.faces[k][0], B.faces[k][1], Ue.data(), Re.data(),
                const_cast<double*>(&B.D[k * size_t(nq2) * size_t(c * 3 * c)]), 2);
      for (int row = 0; row < nd; row++) K[size_t(row) * nd + size_t(col)] = Re[size_t(row)];
    }
  }

  void jac(int e, int q, double J[3][3]) const
  {
    // J[row][col] = jacobians(col,row,q) with jacobians(j,i,q) = Jq[e][j][i][q]
    for (int r = 0; r < 3; r++)
      for (int c = 0; c < 3; c++) J[r][c] = Jq[((size_t(e) * 3 + c) * 3 + r) * nq() + q];
  }

  // One element of evaluation_kernel_impl; store_D selects the AD variant.
  template <int N>
  void element_eval(Integral& I, size_t slot, int e, const double* Ue, double* Re, bool store_D,
                    const double* U1e = nullptr)
  {
    const int c = nc(), NQ = nq();
    std::vector<double> val(size_t(c) * NQ), grd(size_t(c) * 3 * NQ), src(size_t(c) * NQ), flx(size_t(c) * 3 * NQ);
    interpolate(basis, c, Ue, val.data(), grd.data());
    // the second argument's value at the points (its gradient is not used)
    std::vector<double> val1, grd1;
    if (U1e) {
      val1.resize(size_t(c) * NQ), grd1.resize(size_t(c) * 3 * NQ);
      interpolate(basis, c, U1e, val1.data(), grd1.data());
    }
    for (int q = 0; q < NQ; q++) {
      double J[3][3];
      jac(e, q, J);
      // batch_apply_qf (fn/domain_integral_kernels.hpp:125-150): the qfunction
      // works on a copy of the point's state, written back when update_qdata
      double  st[16];
      double* stp = I.ssize ? st : nullptr;
      if (I.ssize) std::copy_n(&I.state[(slot * NQ + q) * I.ssize], I.ssize, st);
      if (!store_D) {
        double v[3], g[9], s[3], f[9];
        for (int i = 0; i < c; i++) {
          v[i] = val[size_t(i) * NQ + q];
          for (int d = 0; d < 3; d++) g[i * 3 + d] = grd[(size_t(i) * 3 + d) * NQ + q];
        }
        parent_to_physical(c, J, g);
        if (U1e) {
          double w[3];
          for (int i = 0; i < c; i++) w[i] = val1[size_t(i) * NQ + q];
          qf_eval2(I.qf, c, v, g, w, s, f, stp);
        } else {
          qf_eval(I.qf, c, v, g, s, f, stp);
        }
        physical_to_parent(c, J, s, f);
        for (int i = 0; i < c; i++) {
          src[size_t(i) * NQ + q] = s[i];
          for (int d = 0; d < 3; d++) flx[(size_t(i) * 3 + d) * NQ + q] = f[i * 3 + d];
        }
      } else {
        // promote_each_to_dual_when (fn/tuple_tensor_dual_functions.hpp:222-272):
        // slot i = value comp i, slot c + 3i + d = parent gradient (i, d)
        dual<N> v[3], g[9], s[3], f[9];
        for (int i = 0; i < c; i++) {
          v[i]             = dual<N>(val[size_t(i) * NQ + q]);
          v[i].gradient[i] = 1.0;
          for (int d = 0; d < 3; d++) {
            g[i * 3 + d]                         = dual<N>(grd[(size_t(i) * 3 + d) * NQ + q]);
            g[i * 3 + d].gradient[c + 3 * i + d] = 1.0;
          }
        }
        parent_to_physical(c, J, g);
        if (U1e) {
          double w[3];
          for (int i = 0; i < c; i++) w[i] = val1[size_t(i) * NQ + q];
          qf_eval2(I.qf, c, v, g, w, s, f, stp);
        } else {
          qf_eval(I.qf, c, v, g, s, f, stp);
        }
        physical_to_parent(c, J, s, f);
        double* D = &I.D[(slot * NQ + q) * I.dsize];
        int     o = 0;
        // get_gradient blocks in the reference's tuple order (:206-213)
        if (I.blk[0][0])
          for (int a = 0; a < c; a++)
            for (int b = 0; b < c; b++) D[o++] = s[a].gradient[b];
        if (I.blk[0][1])
          for (int a = 0; a < c; a++)
            for (int b = 0; b < 3 * c; b++) D[o++] = s[a].gradient[c + b];
        if (I.blk[1][0])
          for (int a = 0; a < 3 * c; a++)
            for (int b = 0; b < c; b++) D[o++] = f[a].gradient[b];
        if (I.blk[1][1])
          for (int a = 0; a < 3 * c; a++)
            for (int b = 0; b < 3 * c; b++) D[o++] = f[a].gradient[c + b];
        for (int i = 0; i < c; i++) {
          src[size_t(i) * NQ + q] = s[i].value;
          for (int d = 0; d < 3; d++) flx[(size_t(i) * 3 + d) * NQ + q] = f[i * 3 + d].value;
        }
      }
      if (I.ssize && update_qdata) std::copy_n(st, I.ssize, &I.state[(slot * NQ + q) * I.ssize]);
    }
    integrate(basis, c, src.data(), flx.data(), Re);
  }

  // chain_rule with the stored derivative (fn/domain_integral_kernels.hpp:223-304)
  void element_jvp(const Integral& I, size_t slot, const double* dUe, double* dRe) const
  {
    const int c = nc(), NQ = nq();
    thread_local std::vector<double> val, grd, src, flx;  // scratch reused across elements
    val.resize(size_t(c) * NQ), grd.resize(size_t(c) * 3 * NQ), src.resize(size_t(c) * NQ);
    flx.resize(size_t(c) * 3 * NQ);
    interpolate(basis, c, dUe, val.data(), grd.data());
    for (int q = 0; q < NQ; q++) {
      const double* D = &I.D[(slot * NQ + q) * I.dsize];
      double        in_v[3], in_g[9], s[3] = {0, 0, 0}, f[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
      for (int i = 0; i < c; i++) {
        in_v[i] = val[size_t(i) * NQ + q];
        for (int d = 0; d < 3; d++) in_g[i * 3 + d] = grd[(size_t(i) * 3 + d) * NQ + q];
      }
      int o = 0;
      if (I.blk[0][0])
        for (int a = 0; a < c; a++)
          for (int b = 0; b < c; b++) s[a] += D[o++] * in_v[b];
      if (I.blk[0][1])
        for (int a = 0; a < c; a++)
          for (int b = 0; b < 3 * c; b++) s[a] += D[o++] * in_g[b];
      if (I.blk[1][0])
        for (int a = 0; a < 3 * c; a++)
          for (int b = 0; b < c; b++) f[a] += D[o++] * in_v[b];
      if (I.blk[1][1])
        for (int a = 0; a < 3 * c; a++)
          for (int b = 0; b < 3 * c; b++) f[a] += D[o++] * in_g[b];
      for (int i = 0; i < c; i++) {
        src[size_t(i) * NQ + q] = s[i];
        for (int d = 0; d < 3; d++) flx[(size_t(i) * 3 + d) * NQ + q] = f[i * 3 + d];
      }
    }
    integrate(basis, c, src.data(), flx.data(), dRe);
  }

  void gather_elem(int e, const double* L, double* Ue) const
  {
    for (int i = 0; i < nc(); i++)
      for (int j = 0; j < space.npe; j++) Ue[size_t(i) * space.npe + j] = L[space.lidx(space.emap[size_t(e) * space.npe + j], i)];
  }
  void scatter_elem(int e, const double* Re, double* L) const
  {
    for (int i = 0; i < nc(); i++)
      for (int j = 0; j < space.npe; j++) L[space.lidx(space.emap[size_t(e) * space.npe + j], i)] += Re[size_t(i) * space.npe + j];
  }

  // residual (store_D = false) or linearize (store_D = true)
  void eval(const double* U, double* R, bool store_D)
  {
    const size_t ndof = size_t(space.ndof());
    std::fill(R, R + ndof, 0.0);
    std::vector<double> Ue(size_t(nc()) * space.npe), Re(size_t(nc()) * space.npe), U1e(size_t(nc()) * space.npe);
    for (auto& I : integrals) {
      if (store_D) I.D.assign(I.elems.size() * size_t(nq()) * I.dsize, 0.0);
      if (I.two && !U1) throw std::runtime_error("a two-argument integral needs its second argument (U1)");
      for (size_t k = 0; k < I.elems.size(); k++) {
        int e = I.elems[k];
        gather_elem(e, U, Ue.data());
        if (I.two) gather_elem(e, U1, U1e.data());
        const double* u1 = I.two ? U1e.data() : nullptr;
        std::fill(Re.begin(), Re.end(), 0.0);
        if (nc() == 1)
          element_eval<4>(I, k, e, Ue.data(), Re.data(), store_D, u1);
        else
          element_eval<12>(I, k, e, Ue.data(), Re.data(), store_D, u1);
        scatter_elem(e, Re.data(), R);
      }
    }
    boundary_eval(U, R, store_D);
  }

  void jvp(const double* dU, double* dR) const
  {
    const size_t ndof = size_t(space.ndof());
    std::fill(dR, dR + ndof, 0.0);
    std::vector<double> Ue(size_t(nc()) * space.npe), Re(size_t(nc()) * space.npe);
    for (auto& I : integrals) {
      if (I.D.empty()) throw std::runtime_error("jvp before linearize");
      for (size_t k = 0; k < I.elems.size(); k++) {
        int e = I.elems[k];
        gather_elem(e, dU, Ue.data());
        std::fill(Re.begin(), Re.end(), 0.0);
        element_jvp(I, k, Ue.data(), Re.data());
        scatter_elem(e, Re.data(), dR);
      }
    }
    boundary_jvp(dU, dR);
  }

  // ---- the threaded CPU baseline ----
  // Thread t takes the contiguous range [lo_t, hi_t) of each integral's
  // elements and scatters into a private buffer that covers only the nodes
  // those elements touch, [nlo_t, nhi_t) (a slab of a lexicographically
  // numbered mesh, so the buffers together hold about one copy of the vector).
  // A final pass sums, per node, the bands that contain it.  The buffers are
  // allocated once per thread count by prepare_threads(), and each thread
  // zeroes and first-touches its own.
  struct Bands {
    int                              nt = 0;
    std::vector<int64_t>             nlo, nhi;
    std::vector<std::vector<double>> buf;  // [t][(node - nlo_t) * nc + c]
  };
  mutable Bands bands;

  static void thread_range(size_t ne, int t, int nt, size_t& lo, size_t& hi)
  {
    lo = ne * size_t(t) / size_t(nt);
    hi = ne * size_t(t + 1) / size_t(nt);
  }

  void prepare_threads(int nthreads) const
  {
    nthreads = std::max(1, nthreads);
    if (bands.nt == nthreads) return;
    Bands b;
    b.nt = nthreads;
    b.nlo.assign(size_t(nthreads), INT64_MAX);
    b.nhi.assign(size_t(nthreads), 0);
    for (int t = 0; t < nthreads; t++) {
      for (auto& I : integrals) {
        size_t lo, hi;
        thread_range(I.elems.size(), t, nthreads, lo, hi);
        for (size_t k = lo; k < hi; k++)
          for (int j = 0; j < space.npe; j++) {
            const int64_t v = space.emap[size_t(I.elems[k]) * space.npe + size_t(j)];
            b.nlo[size_t(t)] = std::min(b.nlo[size_t(t)], v);
            b.nhi[size_t(t)] = std::max(b.nhi[size_t(t)], v + 1);
          }
      }
      if (b.nhi[size_t(t)] <= b.nlo[size_t(t)]) b.nlo[size_t(t)] = b.nhi[size_t(t)] = 0;
    }
    b.buf.resize(size_t(nthreads));
#pragma omp parallel num_threads(nthreads)
    {
#ifdef _OPENMP
      const int t = omp_get_thread_num();
#else
      const int t = 0;
#endif
      if (t < nthreads)  // each thread allocates and first-touches its own band
        b.buf[size_t(t)].assign(size_t(b.nhi[size_t(t)] - b.nlo[size_t(t)]) * size_t(nc()), 0.0);
    }
    bands = std::move(b);
  }

  // out[lidx(node, c)] = sum over the bands holding node, for the nodes of
  // thread t's share (call inside the parallel region, after a barrier)
  void sum_bands(int t, int nt, double* out) const
  {
    const int64_t nn = space.nnodes;
    const int64_t a = nn * t / nt, z = nn * (t + 1) / nt;
    std::vector<int> over;
    for (int s2 = 0; s2 < nt; s2++)
      if (bands.nlo[size_t(s2)] < z && bands.nhi[size_t(s2)] > a) over.push_back(s2);
    for (int64_t i = a; i < z; i++)
      for (int c = 0; c < nc(); c++) {
        double v = 0.0;
        for (int s2 : over)
          if (i >= bands.nlo[size_t(s2)] && i < bands.nhi[size_t(s2)])
            v += bands.buf[size_t(s2)][size_t(i - bands.nlo[size_t(s2)]) * size_t(nc()) + size_t(c)];
        out[space.lidx(i, c)] = v;
      }
  }

  void scatter_band(int e, const double* Re, int t) const
  {
    double*       b  = bands.buf[size_t(t)].data();
    const int64_t lo = bands.nlo[size_t(t)];
    for (int i = 0; i < nc(); i++)
      for (int j = 0; j < space.npe; j++)
        b[size_t(space.emap[size_t(e) * space.npe + j] - lo) * size_t(nc()) + size_t(i)] += Re[size_t(i) * space.npe + j];
  }

  // J.v on `nthreads` host threads (the CPU baseline); jvp() above keeps the
  // reference's serial order.
  void jvp_threads(const double* dU, double* dR, int nthreads) const
  {
    for (auto& I : integrals)
      if (I.D.empty()) throw std::runtime_error("jvp before linearize");
    prepare_threads(nthreads);
    const int nt = bands.nt;
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
      const int tid = omp_get_thread_num();
#else
      const int tid = 0;
#endif
      std::fill(bands.buf[size_t(tid)].begin(), bands.buf[size_t(tid)].end(), 0.0);
      std::vector<double> Ue(size_t(nc()) * space.npe), Re(size_t(nc()) * space.npe);
      for (auto& I : integrals) {
        size_t lo, hi;
        thread_range(I.elems.size(), tid, nt, lo, hi);
        for (size_t k = lo; k < hi; k++) {
          const int e = I.elems[k];
          gather_elem(e, dU, Ue.data());
          std::fill(Re.begin(), Re.end(), 0.0);
          element_jvp(I, k, Ue.data(), Re.data());
          scatter_band(e, Re.data(), tid);
        }
      }
#pragma omp barrier
      sum_bands(tid, nt, dR);
    }
    boundary_jvp(dU, dR);
  }

  // residual / linearization on `nthreads` host threads (sets up the CPU
  // baseline at full size); the stored tangent is the same per element as
  // eval()'s, the residual equal up to the order of the sums.
  void eval_threads(const double* U, double* R, bool store_D, int nthreads)
  {
    prepare_threads(nthreads);
    const int nt = bands.nt;
    for (auto& I : integrals) {
      if (I.two) throw std::runtime_error("eval_threads: two-argument integrals use eval");
      if (store_D) I.D.assign(I.elems.size() * size_t(nq()) * I.dsize, 0.0);
    }
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
      const int tid = omp_get_thread_num();
#else
      const int tid = 0;
#endif
      std::fill(bands.buf[size_t(tid)].begin(), bands.buf[size_t(tid)].end(), 0.0);
      std::vector<double> Ue(size_t(nc()) * space.npe), Re(size_t(nc()) * space.npe);
      for (auto& I : integrals) {
        size_t lo, hi;
        thread_range(I.elems.size(), tid, nt, lo, hi);
        for (size_t k = lo; k < hi; k++) {
          const int e = I.elems[k];
          gather_elem(e, U, Ue.data());
          std::fill(Re.begin(), Re.end(), 0.0);
          if (nc() == 1)
            element_eval<4>(I, k, e, Ue.data(), Re.data(), store_D);
          else
            element_eval<12>(I, k, e, Ue.data(), Re.data(), store_D);
          scatter_band(e, Re.data(), tid);
        }
      }
#pragma omp barrier
      sum_bands(tid, nt, R);
    }
    boundary_eval(U, R, store_D);
  }

  // ---- quantities of interest, Functional<double(trials...)> ----
  // The value (fn/functional_qoi.inl:307-472 with the domain evaluation of
  // fn/domain_integral_kernels.hpp:155-220 at a scalar test space): for each
  // element of the domain, in order, the sum over its points of
  // q(x, u0, u1) detJ w, then the element sums added in element order.
  //   kind 1 measure: q = 1;  2 position: q = x_d (params[0] = d);
  //   3 value: q = u0_c (params[0] = c);  4 Neo-Hookean strain energy of u0:
  //   q = G/2 (tr(F F^T) - 3) - G ln J + lambda/2 (ln J)^2, F = I + grad u0,
  //   lambda = K - 2G/3 (params {density, K, G}); 5 inner: q = u0 . u1.
  double qoi(int kind, const double* params, const std::vector<int>& elems, const double* U0,
             const double* U1) const
  {
    const int        c = nc(), NQ = nq(), Qn = Q;
    std::vector<int> dom = elems;
    if (dom.empty())
      for (int e = 0; e < mesh->ne; e++) dom.push_back(e);
    std::vector<double> Ue(size_t(c) * space.npe), v0(size_t(c) * NQ), g0(size_t(3 * c) * NQ), v1(size_t(c) * NQ),
        g1(size_t(3 * c) * NQ);
    double total = 0.0;
    for (int e : dom) {
      if (U0) {
        gather_elem(e, U0, Ue.data());
        interpolate(basis, c, Ue.data(), v0.data(), g0.data());
      }
      if (U1) {
        gather_elem(e, U1, Ue.data());
        interpolate(basis, c, Ue.data(), v1.data(), g1.data());
      }
      double part = 0.0;
      for (int q = 0; q < NQ; q++) {
        const int qx = q % Qn, qy = (q / Qn) % Qn, qz = q / (Qn * Qn);
        const double w = basis.wq[size_t(qx)] * basis.wq[size_t(qy)] * basis.wq[size_t(qz)];
        double J[3][3];
        jac(e, q, J);
        const double dJ = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                          J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                          J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        double qv = 0.0;
        switch (kind) {
          case 1: qv = 1.0; break;
          case 2: qv = Xq[(size_t(e) * 3 + size_t(params[0])) * NQ + q]; break;
          case 3: qv = v0[size_t(params[0]) * NQ + q]; break;
          case 4: {
            double H[9];
            for (int i = 0; i < 3; i++)
              for (int d = 0; d < 3; d++) H[i * 3 + d] = g0[(size_t(i) * 3 + d) * NQ + q];
            parent_to_physical(3, J, H);
            double F[9], FF = 0.0;
            for (int k = 0; k < 9; k++) {
              F[k] = H[k] + (k % 4 == 0 ? 1.0 : 0.0);
              FF += F[k] * F[k];
            }
            const double detF = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                                F[2] * (F[3] * F[7] - F[4] * F[6]);
            const double K = params[1], G = params[2], lam = K - 2.0 * G / 3.0, lnJ = std::log(detF);
            qv = 0.5 * G * (FF - 3.0) - G * lnJ + 0.5 * lam * lnJ * lnJ;
            break;
          }
          case 5:
            for (int i = 0; i < c; i++) qv += v0[size_t(i) * NQ + q] * v1[size_t(i) * NQ + q];
            break;
          default: throw std::runtime_error("unknown QoI kind");
        }
        part += qv * dJ * w;
      }
      total += part;
    }
    return total;
  }

  // Element gradient K_e[(c,j) test][(c',J) trial], row-major, by applying the
  // stored derivative to each trial shape function (element_gradient_kernel).
  void element_matrix(int integral, size_t slot, double* K) const
  {
    const Integral& I   = integrals[size_t(integral)];
    const int       nd  = nc() * space.npe;
    std::vector<double> Ue(nd, 0.0), Re(nd);
    for (int col = 0; col < nd; col++) {
      std::fill(Ue.begin(), Ue.end(), 0.0);
      Ue[col] = 1.0;
      std::fill(Re.begin(), Re.end(), 0.0);
      element_jvp(I, slot, Ue.data(), Re.data());
      for (int row = 0; row < nd; row++) K[size_t(row) * nd + col] = Re[row];
    }
  }

  void diagonal(double* diag) const
  {
    const size_t ndof = size_t(space.ndof());
    std::fill(diag, diag + ndof, 0.0);
    const int           nd = nc() * space.npe;
    std::vector<double> K(size_t(nd) * nd);
    for (size_t ii = 0; ii < integrals.size(); ii++) {
      const Integral& I = integrals[ii];
      for (size_t k = 0; k < I.elems.size(); k++) {
        element_matrix(int(ii), k, K.data());
        int e = I.elems[k];
        for (int i = 0; i < nc(); i++)
          for (int j = 0; j < space.npe; j++) {
            int r = i * space.npe + j;
            diag[space.lidx(space.emap[size_t(e) * space.npe + j], i)] += K[size_t(r) * nd + r];
          }
      }
    }
    for (size_t bi = 0; bi < bintegrals.size(); bi++) {
      const BIntegral& B = bintegrals[bi];
      for (size_t k = 0; k < B.faces.size(); k++) {
        face_matrix(int(bi), k, K.data());
        int e = B.faces[k][0];
        for (int i = 0; i < nc(); i++)
          for (int j = 0; j < space.npe; j++) {
            int r = i * space.npe + j;
            diag[space.lidx(space.emap[size_t(e) * space.npe + j], i)] += K[size_t(r) * nd + r];
          }
      }
    }
  }
};

// ---------------------------------------------------------------------------
// Solvers.  Essential rows follow physics/solid_mechanics.hpp:1120-1139: the
// residual has essential rows zeroed; the Jacobian is eliminated to identity
// rows and columns.  Jacobi-PCG follows the standard preconditioned CG that
// mfem::CGSolver implements (mfem absent: stopping rule (r, Br) <=
// max(rel^2 (r0, Br0), abs^2); trajectories parity unpinned).
// Newton follows num/equation_solver.cpp:81-235 without line search
// (max_line_search_iterations = 0): norm goal = max(rel * ||r0||, abs),
// x <- x - c, ||.|| = l2.
// ---------------------------------------------------------------------------
inline double dot(const std::vector<double>& a, const std::vector<double>& b)
{
  double s = 0;
  for (size_t i = 0; i < a.size(); i++) s += a[i] * b[i];
  return s;
}

struct SolveReport {
  int    iterations = 0;
  double initial_norm = 0, final_norm = 0;
  bool   converged = false;
  int    line_search_steps = 0;
};

// J_bc v: eliminate columns (v[ess] -> 0), apply J, set rows: out[ess] = v[ess]
inline void apply_bc_jvp(const Form& f, const std::vector<char>& is_ess, const std::vector<double>& v,
                         std::vector<double>& out)
{
  std::vector<double> w = v;
  for (size_t i = 0; i < w.size(); i++)
    if (is_ess[i]) w[i] = 0.0;
  f.jvp(w.data(), out.data());
  for (size_t i = 0; i < w.size(); i++)
    if (is_ess[i]) out[i] = v[i];
}

inline SolveReport pcg_jacobi(const Form& f, const std::vector<char>& is_ess, const std::vector<double>& b,
                              std::vector<double>& x, double rel_tol, double abs_tol, int max_iter)
{
  const size_t        n = b.size();
  std::vector<double> diag(n), r(n), z(n), d(n), Ad(n);
  f.diagonal(diag.data());
  for (size_t i = 0; i < n; i++)
    if (is_ess[i]) diag[i] = 1.0;
  std::fill(x.begin(), x.end(), 0.0);
  r = b;
  for (size_t i = 0; i < n; i++) z[i] = r[i] / diag[i];
  d          = z;
  double nom = dot(z, r);
  SolveReport rep;
  rep.initial_norm = std::sqrt(std::max(nom, 0.0));
  double goal      = std::max(nom * rel_tol * rel_tol, abs_tol * abs_tol);
  // mfem::CGSolver: (r0, B r0) < 0 means B is not positive definite; it stops
  // unconverged with x = 0
  if (nom < 0.0 || nom <= goal) {
    rep.converged  = nom >= 0.0;
    rep.final_norm = rep.initial_norm;
    return rep;
  }
  apply_bc_jvp(f, is_ess, d, Ad);
  double den = dot(d, Ad);
  int    it  = 0;
  for (it = 1; it <= max_iter; it++) {
    double alpha = nom / den;
    for (size_t i = 0; i < n; i++) {
      x[i] += alpha * d[i];
      r[i] -= alpha * Ad[i];
    }
    for (size_t i = 0; i < n; i++) z[i] = r[i] / diag[i];
    double betanom = dot(r, z);
    if (betanom < 0.0) {  // the same test inside the loop: stop unconverged
      rep.converged  = false;
      rep.final_norm = 0.0;
      break;
    }
    if (betanom <= goal) {
      rep.converged  = true;
      rep.final_norm = std::sqrt(betanom);
      break;
    }
    double beta = betanom / nom;
    for (size_t i = 0; i < n; i++) d[i] = z[i] + beta * d[i];
    apply_bc_jvp(f, is_ess, d, Ad);
    den = dot(d, Ad);
    nom = betanom;
    rep.final_norm = std::sqrt(std::max(betanom, 0.0));
  }
  rep.iterations = std::min(it, max_iter);
  return rep;
}

struct NewtonReport {
  int    newton_iterations = 0;
  int    total_linear_iterations = 0;
  double initial_norm = 0, final_norm = 0;
  bool   converged = false;
  int    line_search_steps = 0;
};

// x holds the essential values on entry (and keeps them).
// max_ls: the two-way backtracking line search of serac's NewtonSolver
// (num/equation_solver.cpp:21-185, max_line_search_iterations; 0 is plain
// Newton): halve the step while the residual norm does not fall below the
// previous one; when max_ls halvings fail, search the opposite direction the
// same way; when that fails too, take half the last step of the first
// direction and go on.
inline NewtonReport newton(Form& f, const std::vector<char>& is_ess, std::vector<double>& x, double rel_tol,
                           double abs_tol, int max_iter, double lin_rel, double lin_abs, int lin_max, int max_ls = 0)
{
  const size_t        n = x.size();
  std::vector<double> r(n), c(n);
  auto                residual_norm = [&]() {
    f.eval(x.data(), r.data(), false);
    for (size_t i = 0; i < n; i++)
      if (is_ess[i]) r[i] = 0.0;
    return std::sqrt(dot(r, r));
  };
  NewtonReport rep;
  double       norm = residual_norm();
  rep.initial_norm  = norm;
  double goal       = std::max(rel_tol * norm, abs_tol);
  int    it         = 0;
  for (;; it++) {
    if (norm != norm) break;  // an undefined residual: no convergence
    if (norm <= goal) {
      rep.converged = true;
      break;
    }
    if (it >= max_iter) break;
    f.eval(x.data(), r.data(), true);  // linearize at x (also recomputes r)
    for (size_t i = 0; i < n; i++)
      if (is_ess[i]) r[i] = 0.0;
    SolveReport lr = pcg_jacobi(f, is_ess, r, c, lin_rel, lin_abs, lin_max);
    rep.total_linear_iterations += lr.iterations;
    const double              prev = norm;
    const std::vector<double> x0   = x;
    auto step = [&](double s) {
      for (size_t i = 0; i < n; i++) x[i] = x0[i] + s * c[i];
      return residual_norm();  // a NaN norm never counts as improved
    };
    auto improved = [&](double v) { return v < prev; };
    double scale = 1.0;
    norm         = step(-scale);
    int ls       = 0;
    for (; !improved(norm) && ls < max_ls; ++ls, ++rep.line_search_steps) {
      scale *= 0.5;
      norm = step(-scale);
    }
    if (max_ls > 0 && ls == max_ls && !improved(norm)) {
      scale = 1.0;
      norm  = step(scale);
      ls    = 0;
      for (; !improved(norm) && ls < max_ls; ++ls, ++rep.line_search_steps) {
        scale *= 0.5;
        norm = step(scale);
      }
      if (ls == max_ls && !improved(norm)) {
        ++rep.line_search_steps;
        scale *= 0.5;
        norm = step(-scale);
      }
    }
  }
  rep.newton_iterations = it;
  rep.final_norm        = norm;
  return rep;
}

}  // namespace oracle
