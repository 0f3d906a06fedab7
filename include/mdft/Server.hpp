// Batched and served minimisation over the C ABI (mdft_server_*): many
// solutes of any shape under one HBM budget, same-shape solutes advanced
// together with their functional evaluations fused per tick.  No counterpart
// in the reference, whose Solver runs one field once (src/mini-mdft.hpp:129-133).
#pragma once

#include <array>
#include <complex>
#include <utility>
#include <string>
#include <vector>

#include "Functional.hpp"

namespace MDFT {

struct SolveResult {
  long long id;
  std::array<double, 4> F;  // F, F_id, F_ext, F_exc (kJ/mol)
  int iterations, evaluations;
  int status;  // 0 converged, 1 max_iter, 2 no acceptable step, 3 rejected (budget), 4 non-finite start,
               // 5 set-up failed on admission (error)
  double residual;          // max |h0 g|
  double queued_s, total_s;  // submit -> admission, submit -> end
  std::string error;        // why admission failed (status 5), else empty
};

class Server {
 public:
  // c(q): tuples (np_luc x 5: m, n, mu, nu, khi) and values (np_luc x nq) for
  // every mmax that will be requested, as read by the Luc reader
  Server(const std::vector<int>& tuples, const std::vector<std::complex<double>>& c, int nq, double dq, int mrso,
         double n0, double temperature, Precision precision = Precision::MixedX6, int capacity = 4, int m = 5,
         int max_iter = 400, double gtol = 1e-5, double budget_bytes = 200e9) {
    const int np_luc = static_cast<int>(tuples.size() / 5);
    Impl::Throw_If(np_luc <= 0 || c.size() != static_cast<size_t>(np_luc) * nq, "server: bad c(q) table");
    Impl::check(mdft_server_create(np_luc, tuples.data(), reinterpret_cast<const double*>(c.data()), nq, dq, mrso, n0,
                                   temperature, precision_code<double>(precision), capacity, m, max_iter, gtol,
                                   budget_bytes, &m_h));
  }
  ~Server() {
    if (m_h) (void)mdft_server_destroy(m_h);
  }
  Server(const Server&) = delete;
  Server& operator=(const Server&) = delete;

  // start points of the solves (before the first step): xi = 1 (default) or
  // xi = exp(-V / (2 kT)) from each solute's V_ext
  void start_from_vext(bool on) { Impl::check(mdft_server_start_from_vext(m_h, on ? 1 : 0)); }

  // queues one solute; out_xi (device, N no doubles) receives the final xi
  template <typename ExecutionSpace, typename ScalarType>
  long long submit(const SpatialGrid<ExecutionSpace, ScalarType>& grid, int mmax, const std::vector<Site>& solute,
                   const std::vector<Site>& solvent, double vmax, double* out_xi = nullptr) {
    long long id = 0;
    Impl::check(mdft_server_submit(m_h, grid.m_nx, grid.m_ny, grid.m_nz, grid.m_lx, grid.m_ly, grid.m_lz, mmax,
                                   static_cast<int>(solute.size()), reinterpret_cast<const double*>(solute.data()),
                                   static_cast<int>(solvent.size()), reinterpret_cast<const double*>(solvent.data()),
                                   vmax, out_xi, &id));
    return id;
  }
  // admissions, then one tick of every engine; returns (active, pending)
  std::pair<int, int> step() {
    int a = 0, p = 0;
    Impl::check(mdft_server_step(m_h, &a, &p));
    return {a, p};
  }
  bool poll(SolveResult& r) {
    int got = 0;
    Impl::check(mdft_server_poll(m_h, &r.id, r.F.data(), &r.iterations, &r.evaluations, &r.status, &r.residual,
                                 &r.queued_s, &r.total_s, &got));
    if (got) {
      char buf[1024];
      Impl::check(mdft_server_poll_error(m_h, buf, static_cast<int>(sizeof(buf))));
      r.error = buf;
    }
    return got != 0;
  }
  // steps until nothing is active or waiting; the finished solves in order of completion
  std::vector<SolveResult> run() {
    std::vector<SolveResult> out;
    SolveResult r{};
    for (;;) {
      const auto ap = step();
      while (poll(r)) out.push_back(r);
      if (ap.first == 0 && ap.second == 0) break;
    }
    while (poll(r)) out.push_back(r);
    return out;
  }

 private:
  mdft_server_t m_h = nullptr;
};

}  // namespace MDFT
