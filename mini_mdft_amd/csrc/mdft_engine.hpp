// Batched and served minimisation: many solutes of one shape minimised
// together, one trial point per solute per tick, the functional evaluations
// of a tick fused into one pass of every stage (see mdft_engine.hip).
#pragma once

#include <deque>
#include <map>
#include <memory>
#include <string>
#include <tuple>

#include "mdft_functional.hpp"

namespace mdft {

// One shape (grid, box, angular grid, precision form): capacity slots of
// solutes minimised in lock step.  Per tick every active slot makes one
// trial (or evaluates its start point); the trial points of all slots sit
// back to back in one field so that the excess evaluation of the tick is a
// single batched pass; the L-BFGS passes of all slots are one launch each.
// The kernels of a tick are captured once per batch width as a HIP graph and
// replayed; per-slot scalars travel in a device table uploaded each tick.
class Engine {
 public:
  // nb fields back to back in x / df (double, or float with fp32_fields);
  // ff of field j into ff[j ffstride]
  using ExcessBatch = std::function<void(int nb, const void* x, void* df, double* ff, long ffstride,
                                         hipStream_t s)>;

  // fp32_fields (needs fp32_pairs): the excess sees FP32 copies of the FP64
  // trial points and returns FP32 df
  Engine(const FunctionalParams& p, const std::vector<double>& w, int capacity, const LbfgsOptions& o,
         bool fp32_pairs, ExcessBatch excess, bool fp32_fields = false);
  ~Engine();
  Engine(const Engine&) = delete;
  Engine& operator=(const Engine&) = delete;

  int capacity() const { return capacity_; }
  int active() const;
  bool has_free_slot() const { return active() < capacity_; }
  hipStream_t stream() const { return stream_; }
  // device bytes a solute adds, and those the engine holds for its slots
  size_t job_bytes() const;
  size_t shared_bytes() const { return shared_bytes_; }

  // Takes a solute into a free slot and returns the slot: its V_ext buffer
  // (vext_of) is the caller's to fill on stream(), its start point is xi = 1.
  // When out_xi (device, N no doubles) is given, the final point is copied
  // there when the solve ends.
  int admit(long long id, double* out_xi = nullptr);
  // gives an admitted slot back before its first tick (its set-up failed)
  void cancel(int slot);
  double* vext_of(int slot);
  // called once the slot's V_ext is in place (on stream()): with
  // start_from_vext on, the slot starts from xi = exp(-V / (2 kT))
  // (xi_from_vext) instead of xi = 1
  void vext_ready(int slot);
  // one trial (or start-point evaluation) for every active slot; returns the
  // number of slots that finished in this tick.  begin_tick enqueues the
  // tick on stream(), end_tick waits for it and makes the decisions.
  int tick();
  void begin_tick();
  int end_tick();
  struct Done {
    long long id;
    LbfgsResult r;
  };
  // finished solves since the last call (their slots are free again)
  std::vector<Done> take_done();
  // graphs on (default) or off: results are the same either way
  void use_graphs(bool on) { graphs_ = on; }
  // start points: xi = 1 (default, Solvent::init_density) or exp(-V / (2 kT))
  void start_from_vext(bool on) { start_vext_ = on; }

 private:
  struct Slot;
  void enqueue(int nbatch);
  void release_slot(int j);
  void release();
  int width_ = 0;  // slots of the tick in flight

  FunctionalParams p_;
  LbfgsOptions o_;
  ExcessBatch excess_;
  int capacity_;
  bool fp32_;
  bool f32f_;
  bool graphs_ = true;
  bool start_vext_ = false;
  long n_ = 0;
  int nb_ = 0, step_ = 0;
  hipStream_t stream_ = nullptr;
  double* tab_ = nullptr;    // w | h0 (device)
  double* xslab_ = nullptr;  // capacity trial points, back to back
  float* xfslab_ = nullptr;  // their FP32 copies (fp32_fields)
  void* dfslab_ = nullptr;   // double, or float with fp32_fields
  void* zeros_ = nullptr;    // one zero field of pair values (unused pair columns)
  double* part_ = nullptr;   // capacity x rows x nb partial sums
  double* res_ = nullptr;    // capacity x results
  double* host_res_ = nullptr;
  void* desc_ = nullptr;     // capacity slot descriptors (device)
  void* host_desc_ = nullptr;
  size_t desc_bytes_ = 0, shared_bytes_ = 0;
  std::vector<std::unique_ptr<Slot>> slots_;
  std::vector<Done> done_;
  std::map<int, hipGraphExec_t> graphs_by_width_;
  std::map<int, int> eager_runs_;
};

// Many shapes under one HBM budget: requests wait in a queue and are
// admitted (first fit, in arrival order) while their footprint fits; one
// engine per shape, created on demand and dropped when idle and memory is
// needed; the engines' ticks run concurrently on their own streams.
class Server {
 public:
  struct Shape {
    int nx, ny, nz;
    double lx, ly, lz;
    int mmax;
    bool operator<(const Shape& o) const {
      return std::tie(nx, ny, nz, lx, ly, lz, mmax) < std::tie(o.nx, o.ny, o.nz, o.lx, o.ly, o.lz, o.mmax);
    }
  };
  struct Request {
    long long id;
    Shape shape;
    std::vector<double> solute, solvent;  // rows of 6
    double vmax;
    double* out_xi;  // device buffer for the final point, or null
    double t_submit;
  };
  struct Result {
    long long id;
    LbfgsResult r;
    double queued_s, total_s;  // submit -> admit, submit -> done
    bool rejected;
    std::string error;  // why admission failed (status 5), else empty
  };
  // builds the engines of a shape: (shape) -> engine and the bytes it holds
  // besides its slots (tables, plans, work buffers)
  using EngineFactory = std::function<std::unique_ptr<Engine>(const Shape&, size_t& table_bytes)>;
  // predicts (job bytes, engine bytes) of a shape without building it
  using Footprint = std::function<std::pair<size_t, size_t>(const Shape&)>;

  Server(EngineFactory make, Footprint footprint, double budget_bytes, int mrso);
  long long submit(const Shape& shape, const double* solute, int nsolute, const double* solvent, int nsolvent,
                   double vmax, double* out_xi = nullptr);
  // admits what fits, then one tick of every engine with active solutes
  void step();
  bool poll(Result& out);
  int active() const;
  int pending() const { return static_cast<int>(queue_.size()); }
  double bytes_in_use() const { return in_use_; }
  int engines() const { return static_cast<int>(engines_.size()); }

 private:
  struct Live {
    std::unique_ptr<Engine> engine;
    size_t table_bytes = 0;
    std::map<long long, double> admitted_at;  // id -> time
  };
  void admit();
  EngineFactory make_;
  Footprint footprint_;
  double budget_, in_use_ = 0.0;
  int mrso_;
  long long next_id_ = 1;
  std::deque<Request> queue_;
  std::map<Shape, Live> engines_;
  std::map<long long, Request> running_;
  std::deque<Result> results_;
};

double now_seconds();

}  // namespace mdft
