// batch.hpp -- many mesh solves in one pass: realisations of equal size are stacked into batched
// level arrays and advanced together, one kernel launch per stage for the whole stack, with one
// flexible-PCG iteration replayed as a captured HIP graph.  Included by mmh.hip (host side).
//
// Each realisation follows exactly the arithmetic of the single solve (mmh_solve_d with the
// Krylov option): same kernels, same per-grid reduction order, per-grid CG scalars computed on the
// device.  A realisation that reaches the tolerance gets alpha = beta = 0 from then on, which
// freezes its iterate while the others continue, so every result is bit-identical to its single
// solve.  The run100multigrid.sh workflow of the reference (one process per realisation, serial)
// becomes one call.

// Level arrays of B realisations of size n, stacked per level: field (l, s) at base[l] + s * m^3.
struct BatchGroup {
  int n = 0, B = 0, nlev = 0;
  std::vector<int> items;  // indices into the caller's list
  int m[MMH_MAXLEV];
  double dx[MMH_MAXLEV];
  double* u[MMH_MAXLEV];
  double* rhs[MMH_MAXLEV];
  double* r[MMH_MAXLEV];
  double* G[MMH_MAXLEV];
  double* A[MMH_MAXLEV] = {};  // Galerkin operators of levels 1.. (27 m^3 per grid), if built
  // PCG vectors at the finest level (sub-lattice layout), B n^3 each
  double *x, *b, *cr, *ro, *p, *q, *z;
  // scratch of the per-grid reductions, and per-grid scalars
  double *part, *part2, *bm, *bs, *bs2;
  int* hflags = nullptr;  // pinned host copy of active
  int* act0 = nullptr;    // active at the start of the current iteration: the grids its kernels work on
  double* ssum = nullptr;  // set-up sums: det sums, rhs sums, max|rhs| (B each)
  double *sums, *rz, *pq, *rmax, *rden, *rznew, *zro, *alpha, *beta, *relres;
  int *active, *iters;
  double* pscal = nullptr;  // one-workgroup-per-grid cycles: their level scalars (3 CLEV_MAX per grid)
  std::string tag;
};

void batch_reduce(mmh_ctx* c, BatchGroup& g, const double* x, i64 n, int op, double* dst, const int* act = nullptr) {
  // per grid: the staging of reduce() (blocks of RB*RPT, halving tree), grids side by side
  const i64 ch = (i64)RB * RPT;
  i64 nb = (n + ch - 1) / ch;
  const unsigned B = g.B;
  if (nb <= 1) {
    if (op == 0) k_reduce_stage<double, 0><<<dim3(1, B), RB, 0, c->stream>>>(dst, x, n, n, act);
    else k_reduce_stage<double, 1><<<dim3(1, B), RB, 0, c->stream>>>(dst, x, n, n, act);
    LAUNCHCHK();
    return;
  }
  double* a = g.part;
  double* bb = g.part2;
  if (op == 0) k_reduce_stage<double, 0><<<dim3((unsigned)nb, B), RB, 0, c->stream>>>(a, x, n, n, act);
  else k_reduce_stage<double, 1><<<dim3((unsigned)nb, B), RB, 0, c->stream>>>(a, x, n, n, act);
  LAUNCHCHK();
  while (nb > 1) {
    const i64 nb2 = (nb + ch - 1) / ch;
    double* out = nb2 == 1 ? dst : bb;
    if (op == 0) k_reduce_stage<double, 0><<<dim3((unsigned)nb2, B), RB, 0, c->stream>>>(out, a, nb, nb, act);
    else k_reduce_stage<double, 1><<<dim3((unsigned)nb2, B), RB, 0, c->stream>>>(out, a, nb, nb, act);
    LAUNCHCHK();
    std::swap(a, bb);
    nb = nb2;
  }
}

std::pair<int, int> batch_shape(mmh_ctx* c, int m, int B) {
  auto it = c->tune.find(std::to_string(m) + ":" + std::to_string(B));
  return it == c->tune.end() ? std::make_pair(256, 1) : it->second;
}

void batch_sweeps(mmh_ctx* c, BatchGroup& g, double* phi, const double* rhs, const double* G, int l, int ns, bool reverse,
                  std::pair<int, int> shape, bool from_zero = false, const int* act = nullptr) {
  const int m = g.m[l];
  const i64 h3 = cube(m / 2);
  const double dx2 = g.dx[l] * g.dx[l];
  const int bsz = shape.first;
  const dim3 grid(nblk(h3, bsz), g.B);
  for (int sw = 0; sw < ns; ++sw)
    for (int cc = 0, done = 0; cc < 8; ++cc) {
      const int col = reverse ? 7 - cc : cc;
      if (sw == 0 && from_zero)
        k_colour<double, 1, false, true><<<grid, bsz, 0, c->stream>>>(phi, rhs, G, m, dx2, col >> 2, (col >> 1) & 1,
                                                                      col & 1, cube(m), shape.second, (unsigned)done,
                                                                      act);
      else
        k_colour<double, 1><<<grid, bsz, 0, c->stream>>>(phi, rhs, G, m, dx2, col >> 2, (col >> 1) & 1, col & 1, cube(m),
                                                         shape.second, 0xFFu, act);
      done |= 1 << col;
    }
  LAUNCHCHK();
}

// Sweeps of a batched Galerkin level (gal_sweeps for the stack).
void batch_gal_sweeps(mmh_ctx* c, BatchGroup& g, double* phi, const double* rhs, int l, int ns, bool reverse,
                      bool from_zero = false, const int* act = nullptr) {
  const int m = g.m[l];
  const dim3 grid(nblk(cube(m / 2)), g.B);
  for (int sw = 0; sw < ns; ++sw)
    for (int cc = 0, done = 0; cc < 8; ++cc) {
      const int col = reverse ? 7 - cc : cc;
      if (sw == 0 && from_zero)
        k_gal_colour<double, true><<<grid, 256, 0, c->stream>>>(phi, rhs, g.A[l], m, col >> 2, (col >> 1) & 1, col & 1,
                                                                cube(m), (unsigned)done, act);
      else
        k_gal_colour<double><<<grid, 256, 0, c->stream>>>(phi, rhs, g.A[l], m, col >> 2, (col >> 1) & 1, col & 1,
                                                          cube(m), 0xFFu, act);
      done |= 1 << col;
    }
  LAUNCHCHK();
}

// Symmetric V(nu, nu) cycle of the batch from level l, the arithmetic of vcycle<double> with the
// Krylov schedule (no guard, reverse colour order on the way up, mean removal on every level).
// act: grids y with act[y] = 0 are left alone (null: every grid).
// Levels l.. of the group as a stack table for the one-workgroup-per-grid kernels (level l's u and
// rhs are u0 and rhs0 when l = 0).
StackLevels stack_levels(const BatchGroup& g, int l, double* u0, double* rhs0, const Sched& S) {
  StackLevels L{};
  L.nl = g.nlev - l;
  for (int q = 0; q < L.nl; ++q) {
    const int k = l + q;
    L.u[q] = k == 0 ? u0 : g.u[k];
    L.rhs[q] = k == 0 ? rhs0 : g.rhs[k];
    L.r[q] = g.r[k];
    L.G[q] = g.G[k];
    L.A[q] = g.A[k];
    L.m[q] = g.m[k];
    L.sw[q] = S.pre(g.m[k]);
    L.dx[q] = g.dx[k];
  }
  L.coarse = S.coarse();
  L.symmetric = S.symmetric ? 1 : 0;
  L.scal = g.pscal;
  return L;
}

// Whether levels l.. run as one workgroup per grid (k_coarse_stack): levels up to the size set by
// option "batch_coarse_fused", with a schedule that sweeps as often up as down on each of them.
bool batch_coarse_applies(const mmh_ctx* c, const BatchGroup& g, int l, const Sched& S) {
  if (!c->batch_coarse_fused || g.m[l] > std::min(c->batch_coarse_fused, CTOP_MAX) || g.nlev - l > CLEV_MAX) return false;
  for (int k = l; k < g.nlev; ++k)
    if (S.pre(g.m[k]) != S.post(g.m[k])) return false;
  return true;
}

void batch_vcycle(mmh_ctx* c, BatchGroup& g, int l, double* u0, double* rhs0, const Sched& S, const int* act = nullptr) {
  if (batch_coarse_applies(c, g, l, S)) {
    k_coarse_stack<<<g.B, CW, 0, c->stream>>>(stack_levels(g, l, u0, rhs0, S), act);
    LAUNCHCHK();
    return;
  }
  const int m = g.m[l];
  const i64 m3 = cube(m);
  double* u = l == 0 ? u0 : g.u[l];
  double* rh = l == 0 ? rhs0 : g.rhs[l];
  const auto shape = batch_shape(c, m, g.B);
  if (m == 4) {
    if (g.A[l]) {
      batch_gal_sweeps(c, g, u, rh, l, S.coarse(), false, c->zero_start, act);
      batch_gal_sweeps(c, g, u, rh, l, S.coarse(), S.symmetric, false, act);
    } else {
      batch_sweeps(c, g, u, rh, g.G[l], l, S.coarse(), false, shape, c->zero_start, act);
      batch_sweeps(c, g, u, rh, g.G[l], l, S.coarse(), S.symmetric, shape, false, act);
    }
    batch_reduce(c, g, u, m3, 0, g.sums, act);
    k_zerosum_guard<double><<<dim3(1, g.B), 256, 0, c->stream>>>(u, g.sums, nullptr, m3, 0, nullptr, 1, act);
    LAUNCHCHK();
    return;
  }
  const int pre = S.pre(m);
  if (g.A[l]) {
    batch_gal_sweeps(c, g, u, rh, l, pre, false, c->zero_start, act);
    k_gal_resid<double><<<dim3(gal_resid_blocks(m), g.B), 256, 0, c->stream>>>(g.r[l], u, rh, g.A[l], m, m3, act);
    LAUNCHCHK();
  } else {
    batch_sweeps(c, g, u, rh, g.G[l], l, pre, false, shape, c->zero_start, act);
    op_resid<double>(c, g.r[l], g.bm, u, rh, g.G[l], m, 1.0 / (g.dx[l] * g.dx[l]), false, g.B, 1, act);
  }
  const int nrmax = std::min(4, std::max(1, m / 4));
  const int nrl = pre > nrmax ? (pre % nrmax ? pre % nrmax : nrmax) : pre;
  k_restrict_resid<double, 1><<<dim3(nblk(cube(m / 2)), g.B), 256, 0, c->stream>>>(g.rhs[l + 1], g.r[l], m,
                                                                                   nrl % (m / 2), 1, act);
  k_fill<double><<<dim3(nblk(cube(m / 2)), g.B), 256, 0, c->stream>>>(g.u[l + 1], 0.0, cube(m / 2), cube(m / 2), act);
  LAUNCHCHK();
  batch_vcycle(c, g, l + 1, u0, rhs0, S, act);
  k_prolong<double, 1><<<dim3(nblk(cube(m / 2)), g.B), 256, 0, c->stream>>>(u, g.u[l + 1], m, 1, 1, act);
  LAUNCHCHK();
  if (g.A[l]) batch_gal_sweeps(c, g, u, rh, l, S.post(m), S.symmetric, false, act);
  else batch_sweeps(c, g, u, rh, g.G[l], l, S.post(m), S.symmetric, shape, false, act);
  batch_reduce(c, g, u, m3, 0, g.sums, act);
  k_zerosum_guard<double><<<dim3(nblk(m3), g.B), 256, 0, c->stream>>>(u, g.sums, nullptr, m3, 0, nullptr, 1, act);
  LAUNCHCHK();
}

// z.a and z.b of every grid in one pass over z (k_dot2, as the single solve's dot products).
void batch_dot2(mmh_ctx* c, BatchGroup& g, const double* z, const double* a, const double* b, double* dst_a,
                double* dst_b, const int* act = nullptr) {
  const i64 n3 = cube(g.n);
  const unsigned nd = nblk(n3, RB * RPT);
  k_dot2<double><<<dim3(nd, g.B), RB, 0, c->stream>>>(g.bs, g.bs2, z, a, b, n3, nullptr, n3, act);
  LAUNCHCHK();
  batch_reduce(c, g, g.bs, nd, 0, dst_a, act);
  batch_reduce(c, g, g.bs2, nd, 0, dst_b, act);
}

void batch_precond(mmh_ctx* c, BatchGroup& g, const Sched& S, const int* act = nullptr) {
  const i64 n3 = cube(g.n);
  k_fill<double><<<dim3(nblk(n3), g.B), 256, 0, c->stream>>>(g.z, 0.0, n3, n3, act);
  LAUNCHCHK();
  batch_vcycle(c, g, 0, g.z, g.cr, S, act);
}

// One flexible-PCG iteration for every grid of the batch (the loop body of pcg()).  The grids
// that had converged before it (act0[y] = 0, set at the alpha step) are skipped by every kernel:
// their alpha = beta = 0 would leave x and r unchanged, and nothing reads their other vectors
// again.  The operator runs before this iteration's alpha step, on the previous iteration's
// flags (a grid that converged then is computed once more, unused).
void batch_iteration(mmh_ctx* c, BatchGroup& g, const Sched& S, double tol) {
  const i64 n3 = cube(g.n);
  const unsigned nb = nblk(n3, RB), sb = nblk(g.B, 64);
  const int* act = g.act0;
  const unsigned ns = op_apply<double>(c, g.q, g.bs, g.p, g.G[0], g.n, 1.0, g.B, 1, act);
  batch_reduce(c, g, g.bs, ns, 0, g.pq, act);
  k_cg_scalars<<<sb, 64, 0, c->stream>>>(0, g.B, 0, tol, g.alpha, g.beta, g.rz, g.pq, g.rmax, g.rden, g.rznew, g.zro,
                                         g.active, g.iters, g.relres, g.act0);
  k_cg_update_b<<<dim3(nb, g.B), RB, 0, c->stream>>>(g.cr, g.ro, g.q, g.alpha, g.bm, n3, act);
  LAUNCHCHK();
  batch_reduce(c, g, g.bm, nb, 1, g.rmax, act);
  k_cg_scalars<<<sb, 64, 0, c->stream>>>(1, g.B, 0, tol, g.alpha, g.beta, g.rz, g.pq, g.rmax, g.rden, g.rznew, g.zro,
                                         g.active, g.iters, g.relres, g.act0);
  batch_precond(c, g, S, act);
  batch_dot2(c, g, g.z, g.cr, g.ro, g.rznew, g.zro, act);
  k_cg_scalars<<<sb, 64, 0, c->stream>>>(2, g.B, 0, tol, g.alpha, g.beta, g.rz, g.pq, g.rmax, g.rden, g.rznew, g.zro,
                                         g.active, g.iters, g.relres, g.act0);
  k_cg_dir_b<<<dim3(nblk(n3), g.B), 256, 0, c->stream>>>(g.x, g.p, g.z, g.alpha, g.beta, n3, act);
  LAUNCHCHK();
}

// Chooses the colour kernel's block size and XCD remap per level size by timing batched sweeps.
void batch_autotune(mmh_ctx* c, BatchGroup& g) {
  const int cand_b[3] = {64, 128, 256};
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  for (int l = 0; l < g.nlev; ++l) {
    const int m = g.m[l];
    const std::string key = std::to_string(m) + ":" + std::to_string(g.B);
    if (c->tune.count(key)) continue;
    if (m < 16) { c->tune[key] = {64, 0}; continue; }
    double* phi = l == 0 ? g.z : g.u[l];
    const double* rh = l == 0 ? g.cr : g.rhs[l];
    float best = 1e30f;
    std::pair<int, int> pick{256, 1};
    for (int bi = 0; bi < 3; ++bi)
      for (int rm = 0; rm < 2; ++rm) {
        const std::pair<int, int> sh{cand_b[bi], rm};
        batch_sweeps(c, g, phi, rh, g.G[l], l, 1, false, sh);  // warm
        HIPCHK(hipEventRecord(e0, c->stream));
        batch_sweeps(c, g, phi, rh, g.G[l], l, 2, false, sh);
        HIPCHK(hipEventRecord(e1, c->stream));
        HIPCHK(hipEventSynchronize(e1));
        float ms = 0;
        HIPCHK(hipEventElapsedTime(&ms, e0, e1));
        if (ms < best) { best = ms; pick = sh; }
      }
    c->tune[key] = pick;
  }
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
}

void batch_alloc(mmh_ctx* c, BatchGroup& g) {
  const std::string t = "bg" + std::to_string(g.n) + "x" + std::to_string(g.B) + "_";
  g.tag = t;
  int nl = 0;
  for (int m = g.n; m >= 4; m >>= 1) ++nl;
  g.nlev = nl;
  const i64 B = g.B;
  for (int l = 0; l < nl; ++l) {
    const int m = g.n >> l;
    g.m[l] = m;
    g.dx[l] = l == 0 ? 1.0 : 2.0 * g.dx[l - 1];
    const std::string tl = t + std::to_string(l);
    g.u[l] = l == 0 ? nullptr : wsT<double>(c, tl + "u", B * cube(m));
    g.rhs[l] = l == 0 ? nullptr : wsT<double>(c, tl + "rhs", B * cube(m));
    g.r[l] = wsT<double>(c, tl + "r", B * cube(m));
    g.G[l] = wsT<double>(c, tl + "G", B * 6 * cube(m));
    g.A[l] = (l > 0 && c->galerkin) ? wsT<double>(c, tl + "A", B * 27 * cube(m)) : nullptr;
  }
  const i64 n3 = cube(g.n);
  g.x = wsT<double>(c, t + "x", B * n3);
  g.pscal = wsT<double>(c, t + "pscal", B * 3 * CLEV_MAX);
  g.b = wsT<double>(c, t + "b", B * n3);
  g.cr = wsT<double>(c, t + "cr", B * n3);
  g.ro = wsT<double>(c, t + "ro", B * n3);
  g.p = wsT<double>(c, t + "p", B * n3);
  g.q = wsT<double>(c, t + "q", B * n3);
  g.z = wsT<double>(c, t + "z", B * n3);
  const i64 nbmax = nblk(n3, RB);
  g.part = wsT<double>(c, t + "part", B * nbmax);
  g.part2 = wsT<double>(c, t + "part2", B * nbmax);
  g.bm = wsT<double>(c, t + "bm", B * nbmax);
  g.bs = wsT<double>(c, t + "bs", B * nbmax);
  g.bs2 = wsT<double>(c, t + "bs2", B * nblk(n3, RB * RPT));
  double* sc = wsT<double>(c, t + "scal", 10 * B);
  g.ssum = wsT<double>(c, t + "setup_sums", 8 * B);
  g.sums = sc; g.rz = sc + B; g.pq = sc + 2 * B; g.rmax = sc + 3 * B; g.rden = sc + 4 * B;
  g.rznew = sc + 5 * B; g.zro = sc + 6 * B; g.alpha = sc + 7 * B; g.beta = sc + 8 * B; g.relres = sc + 9 * B;
  int* ic = (int*)ws_get(c, t + "flags", sizeof(int) * 3 * B);
  g.active = ic; g.iters = ic + B; g.act0 = ic + 2 * B;
}

// Level data of item s: the metric of every level from def and u(1) (build_hier's arithmetic),
// the projected right-hand side and the initial guess in sub-lattice order.  Returns max|rhs|.
// Set-up of every grid of the group, one launch per stage for all of them (grids stacked,
// gridDim.y = B) and all on the device: the caller's arrays are gathered through a table of
// pointers (tab keeps the host copy alive), then per level the restricted deformation and density
// and the metric, the Galerkin operators, rgzerosum (the projected rhs also goes back to the
// caller's array, as in the single solve), max|rhs| into g.ssum[2B + s], and the iterate and
// right-hand side in sub-lattice order.  Per grid this is the single solve's arithmetic (reductions
// by batch_reduce keep reduce()'s staging), so results stay bit-identical to it.
void batch_setup_group(mmh_ctx* c, BatchGroup& g, std::vector<const void*>& tab, double* const* u, double* const* rhs,
                       double* const* defpot, double* const* ubig, int nu) {
  const int n = g.n, B = g.B;
  const i64 n3 = cube(n);
  tab.resize(4 * (size_t)B);
  for (int s = 0; s < B; ++s) {
    const int i = g.items[s];
    tab[s] = ubig[i];
    tab[B + s] = defpot[i];
    tab[2 * B + s] = rhs[i];
    tab[3 * B + s] = u[i];
  }
  const void** dtab = (const void**)ws_get(c, g.tag + "ptrs", sizeof(void*) * 4 * B);
  HIPCHK(hipMemcpyAsync(dtab, tab.data(), sizeof(void*) * 4 * B, hipMemcpyHostToDevice, c->stream));
  const double* const* t_ubig = (const double* const*)dtab;
  const double* const* t_def = (const double* const*)(dtab + B);
  double* const* t_rhs = (double* const*)(dtab + 2 * B);
  const double* const* t_u = (const double* const*)(dtab + 3 * B);
  double* wst = wsT<double>(c, g.tag + "ws0", B * n3);
  double* dst = wsT<double>(c, g.tag + "ds0", B * n3);
  double* rst = wsT<double>(c, g.tag + "rs", B * n3);
  const dim3 gr(nblk(n3), B);
  k_gather<double><<<gr, 256, 0, c->stream>>>(wst, t_ubig, n3, nu);  // density: component 1 of u
  k_gather<double><<<gr, 256, 0, c->stream>>>(dst, t_def, n3, 1);
  k_gather<double><<<gr, 256, 0, c->stream>>>(rst, (const double* const*)t_rhs, n3, 1);
  LAUNCHCHK();
  const double* dl = dst;
  const double* wl = wst;
  for (int l = 0; l < g.nlev; ++l) {
    const int m = g.m[l];
    const dim3 gl(nblk(cube(m)), B);
    if (l > 0) {
      double* d2 = wsT<double>(c, g.tag + "ds" + std::to_string(l), B * cube(m));
      double* w2 = wsT<double>(c, g.tag + "ws" + std::to_string(l), B * cube(m));
      k_restrict27<double><<<gl, 256, 0, c->stream>>>(d2, dl, 2 * m, 1, 1);
      k_restrict27<double><<<gl, 256, 0, c->stream>>>(w2, wl, 2 * m, 1, 1);
      LAUNCHCHK();
      dl = d2;
      wl = w2;
    }
    k_metric<double, 1><<<gl, 256, 0, c->stream>>>(g.G[l], dl, wl, 1, m, g.dx[l], 0, 1);
    LAUNCHCHK();
  }
  // Galerkin operators (build_galerkin's kernels on the stacked levels)
  for (int l = 1; l < g.nlev; ++l) {
    if (!g.A[l]) break;
    const dim3 gl(nblk(cube(g.m[l])), B);
    if (l == 1) k_rap19<double><<<gl, 256, 0, c->stream>>>(g.A[1], g.G[0], n, 1.0 / (g.dx[0] * g.dx[0]), 1);
    else k_rap27<<<gl, 256, 0, c->stream>>>(g.A[l], g.A[l - 1], g.m[l - 1], 1);
    LAUNCHCHK();
  }
  // rgzerosum per grid (mmh/gauss1.f90:1219-1417: det, its sum and the rhs sum, the update), then
  // max|rhs|, the projected rhs back to the caller, the initial iterates and rhs in sub-lattice order
  double* det = wsT<double>(c, g.tag + "det", B * n3);
  double* dsum = g.ssum;
  double* asum = g.ssum + B;
  double* rnmax = g.ssum + 2 * B;
  k_det<double><<<gr, 256, 0, c->stream>>>(det, dst, n, 1.0, 1);
  LAUNCHCHK();
  batch_reduce(c, g, det, n3, 0, dsum);
  batch_reduce(c, g, rst, n3, 0, asum);
  k_rgz_update_b<double><<<gr, 256, 0, c->stream>>>(rst, det, dsum, asum, n3);
  LAUNCHCHK();
  batch_reduce(c, g, rst, n3, 1, rnmax);
  k_scatter<double><<<gr, 256, 0, c->stream>>>(t_rhs, rst, n3);
  k_gather<double><<<gr, 256, 0, c->stream>>>(det, t_u, n3, 1);  // det is free again
  k_relayout<double, 1><<<gr, 256, 0, c->stream>>>(g.x, det, n, 1, 1);
  k_relayout<double, 1><<<gr, 256, 0, c->stream>>>(g.b, rst, n, 1, 1);
  LAUNCHCHK();
}

void solve_batch_dev(mmh_ctx* c, int count, const int* ns, double* const* u, double* const* rhs,
                     double* const* defpot, double* const* ubig, int nu, const mmh_solve_opts* o,
                     mmh_batch_stats* st) {
  require(count >= 1 && ns && u && rhs && defpot && ubig, "solve_batch: bad arguments");
  require(o != nullptr && o->krylov == 1 && o->schedule == 1 && o->nu_pre == o->nu_post && o->mixed == 0,
          "solve_batch: FP64 Krylov solve with a fixed symmetric schedule only");
  require(o->max_cycles >= 1 && o->max_cycles <= 256 && o->tol > 0, "solve_batch: max_cycles in [1, 256], tol > 0");
  require(nu >= 1, "solve_batch: nu >= 1");
  Sched S;
  S.kind = 1; S.nu_pre = o->nu_pre; S.nu_post = o->nu_post; S.nu_coarse = o->nu_coarse;
  S.guard = false; S.symmetric = true; S.nu_top = c->nu_top;
  require(S.nu_pre >= 1 && S.nu_coarse >= 1, "solve_batch: sweep counts >= 1");
  // Graph capture needs a stream other than the null stream: the batch runs on its own stream.
  StreamSwap swap_guard(c);
  std::map<int, std::vector<int>> bysize;
  for (int i = 0; i < count; ++i) {
    require(is_pow2(ns[i]) && ns[i] >= 8 && ns[i] <= 1024, "solve_batch: sizes must be powers of two in [8, 1024]");
    bysize[ns[i]].push_back(i);
  }
  // Phase 1, queued on the solve's stream without host round trips, largest grids first: per size
  // group, set-up, the first residual and preconditioner application, then an event.  Phase 2: the
  // PCG iterations, each group on a stream of its own (option batch_streams) that waits for its
  // group's event only, so the large group iterates while the small ones are still being set up
  // and the small groups' launches overlap its work; one iteration of every unfinished group per
  // round, their active flags copied to pinned host memory and read once per round (a group whose
  // grids all start converged runs one iteration with alpha = beta = 0, which changes nothing).
  // Phase 3: final residuals and results.  A group issues the same launches in the same order
  // either way, so results do not depend on the option.
  struct GroupRun {
    BatchGroup g;
    std::vector<double> rden;
    std::vector<const void*> tab;  // the group's pointer table (host copy)
    hipStream_t s = nullptr;
    int it = 0;
    bool done = false;
    std::unique_ptr<Replay> graph;
    bool persistent = false;  // the whole PCG of each grid in one workgroup (k_pcg_stack)
  };
  struct Streams {  // group streams and pinned flags, released on every exit path
    mmh_ctx* c;
    hipStream_t main;
    std::vector<hipStream_t> own;
    std::vector<hipEvent_t> ev;
    int* hflags = nullptr;
    ~Streams() {
      for (hipStream_t q : own) {
        (void)hipStreamSynchronize(q);
        (void)hipStreamDestroy(q);
      }
      (void)hipStreamSynchronize(main);
      for (hipEvent_t e : ev) (void)hipEventDestroy(e);
      if (hflags) (void)hipHostFree(hflags);
      c->stream = main;
    }
  } streams{c, c->stream, {}, {}, nullptr};
  std::vector<GroupRun> runs(bysize.size());
  int total_items = 0;
  {
    size_t gi = 0;
    for (auto it = bysize.rbegin(); it != bysize.rend(); ++it) {
      auto& kv = *it;
      GroupRun& R = runs[gi++];
      BatchGroup& g = R.g;
      g.n = kv.first;
      g.B = (int)kv.second.size();
      g.items = kv.second;
      total_items += g.B;
      batch_alloc(c, g);
      S.ng = g.n;
      const i64 n3 = cube(g.n);
      R.rden.assign(g.B, 1.0);
      batch_setup_group(c, g, R.tab, u, rhs, defpot, ubig, nu);
      k_batch_flags<<<nblk(g.B, 64), 64, 0, c->stream>>>(g.B, g.ssum + 2 * g.B, g.rden, g.active, g.iters, g.act0);
      k_zero_inactive<<<dim3(nblk(n3), g.B), 256, 0, c->stream>>>(g.x, g.active, n3);
      LAUNCHCHK();
      fill<double>(c, g.relres, 0.0, g.B);
      // the whole solve in one workgroup per grid: grids up to 64^3 (block sums fit the kernel's
      // LDS) and a symmetric schedule (the one-workgroup cycle sweeps as often up as down)
      R.persistent = g.n <= c->batch_persistent && g.n <= 64 && g.nlev <= CLEV_MAX;
      for (int l = 0; l < g.nlev; ++l) R.persistent = R.persistent && S.pre(g.m[l]) == S.post(g.m[l]);
      if (!R.persistent) {
        if (c->batch_autotune) batch_autotune(c, g);
        // r = b - L x, z = M r, rz = <r, z>, p = z
        op_resid<double>(c, g.cr, g.bm, g.x, g.b, g.G[0], g.n, 1.0, false, g.B, 1);
        batch_precond(c, g, S);
        batch_dot2(c, g, g.z, g.cr, g.cr, g.rz, g.zro);  // zro: set again before its first use
        HIPCHK(hipMemcpyAsync(g.p, g.z, sizeof(double) * g.B * n3, hipMemcpyDeviceToDevice, c->stream));
      }
      hipEvent_t e = nullptr;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      streams.ev.push_back(e);
      HIPCHK(hipEventRecord(e, c->stream));
    }
  }
  HIPCHK(hipHostMalloc((void**)&streams.hflags, sizeof(int) * total_items, hipHostMallocDefault));
  {
    int off = 0;
    for (GroupRun& R : runs) {
      if (c->batch_streams) {
        hipStream_t q = nullptr;
        HIPCHK(hipStreamCreateWithFlags(&q, hipStreamNonBlocking));
        streams.own.push_back(q);
        R.s = q;
        HIPCHK(hipStreamWaitEvent(q, streams.ev[&R - runs.data()], 0));  // the group's set-up
      } else {
        R.s = streams.main;
      }
      R.g.hflags = streams.hflags + off;
      off += R.g.B;
    }
  }
  auto copy_flags = [&](GroupRun& R) {
    HIPCHK(hipMemcpyAsync(R.g.hflags, R.g.active, sizeof(int) * R.g.B, hipMemcpyDeviceToHost, R.s));
  };
  auto any_active = [&](const GroupRun& R) {
    for (int s = 0; s < R.g.B; ++s)
      if (R.g.hflags[s]) return true;
    return false;
  };
  for (bool pending = true; pending;) {
    pending = false;
    for (GroupRun& R : runs) {
      if (R.done) continue;
      c->stream = R.s;
      S.ng = R.g.n;
      if (R.persistent) {
        // the whole solve of every grid, one workgroup each
        BatchGroup& g = R.g;
        PcgStack P{};
        P.n = g.n; P.B = g.B; P.max_it = o->max_cycles; P.tol = o->tol;
        P.L = stack_levels(g, 0, g.z, g.cr, S);
        P.x = g.x; P.b = g.b; P.cr = g.cr; P.ro = g.ro; P.p = g.p; P.q = g.q; P.z = g.z;
        P.rden = g.rden; P.active = g.active; P.iters = g.iters; P.relres = g.relres;
        k_pcg_stack<<<g.B, CW, 0, c->stream>>>(P);
        LAUNCHCHK();
        R.it = o->max_cycles;
        copy_flags(R);
        pending = true;
        continue;
      }
      // one iteration captured as a graph (batch_alloc sized every buffer it uses) and replayed
      if (c->batch_graph) {
        if (!R.graph) {
          R.graph.reset(new Replay(c));
          R.graph->capture([&] { batch_iteration(c, R.g, S, o->tol); });
        }
        R.graph->launch();
      } else {
        batch_iteration(c, R.g, S, o->tol);
      }
      ++R.it;
      copy_flags(R);
      pending = true;
    }
    for (GroupRun& R : runs) {
      if (R.done) continue;
      HIPCHK(hipStreamSynchronize(R.s));
      if (!any_active(R) || R.it >= o->max_cycles) R.done = true;
    }
  }
  bool all_ok = true;
  for (GroupRun& R : runs) {
    BatchGroup& g = R.g;
    c->stream = R.s;
    const i64 n3 = cube(g.n);
    const bool group_ok = !any_active(R);
    all_ok = all_ok && group_ok;
    // true residual of the final iterates, mean removal, back to the caller's layout
    const unsigned nsr = op_resid<double>(c, (double*)nullptr, g.bm, g.x, g.b, g.G[0], g.n, 1.0, false, g.B, 1);
    batch_reduce(c, g, g.bm, nsr, 1, g.rmax);
    batch_reduce(c, g, g.x, n3, 0, g.sums);
    k_zerosum_guard<double><<<dim3(nblk(n3), g.B), 256, 0, c->stream>>>(g.x, g.sums, nullptr, n3, 0, nullptr, 1);
    LAUNCHCHK();
    std::vector<double> hr(g.B);
    std::vector<int> hi(g.B);
    HIPCHK(hipMemcpyAsync(hr.data(), g.rmax, sizeof(double) * g.B, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(hi.data(), g.iters, sizeof(int) * g.B, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipMemcpyAsync(R.rden.data(), g.rden, sizeof(double) * g.B, hipMemcpyDeviceToHost, c->stream));
    for (int s = 0; s < g.B; ++s) relayout(c, u[g.items[s]], g.x + (i64)s * n3, g.n, 1, 0);
    HIPCHK(hipStreamSynchronize(c->stream));
    if (st)
      for (int s = 0; s < g.B; ++s) {
        st[g.items[s]].cycles = hi[s];
        st[g.items[s]].relres = hr[s] / R.rden[s];
        st[g.items[s]].converged = g.hflags[s] ? 0 : 1;
      }
  }
  runs.clear();  // graphs go before their streams
  c->stream = streams.main;
  if (!all_ok) throw MmhError(MMH_ENOCONV, "solve_batch: tolerance not reached for every realisation");
}
