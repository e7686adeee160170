// Request server: continuous batching of adapt() requests on one GPU.
//
// Each slot is a host worker thread with its own context (HIP stream and
// allocator), so the passes of concurrent requests overlap on the device.
// Requests are admitted in arrival order when a slot is free and their
// predicted HBM footprint fits in the budget next to the running ones (a
// request larger than the whole budget runs alone).  Mesh state stays on
// the device for the whole request: the box is built there, adapted pass by
// pass, and only counts, timings and an optional digest come back.
//
// The footprint prediction uses the metric complexity of the box
// (get_complexity / get_expected_nelems, src/Omega_h_metric.cpp:464-499, with
// the typical unit simplex sizes of :288-290).
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>

#include "../../include/omega_h_amd.h"
#include "osh_guard.hpp"
#include "osh_passes.hpp"

namespace osh {

std::uint64_t mesh_digest(Ctx& c, DMesh& m);
void to_options(osh_adapt_opts const* o, AdaptOptions& r);

namespace {

double wall() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

constexpr double kBytesPerElement = 1024.0;

double typical_unit_simplex_size(int dim) {
  return dim == 3 ? 0.0838934100219 : (dim == 2 ? 0.377645136635 : 1.0);
}

double sqrt_det_of_constant_metric(int dim, int ncomps, double const* m) {
  if (ncomps == 1) return std::pow(m[0], 0.5 * dim);
  if (dim == 2) return std::sqrt(m[0] * m[1] - m[2] * m[2]);
  // xx, yy, zz, xy, yz, xz
  double a = m[0], b = m[3], c = m[5], e = m[1], f = m[4], i = m[2];
  double det = a * (e * i - f * f) - b * (b * i - f * c) + c * (b * f - e * c);
  return std::sqrt(det > 0 ? det : 0.0);
}

struct Request {
  int ticket;
  int dim, nx, ny, nz;
  int ncomps;
  double metric[6];
  AdaptOptions opts;
  bool want_digest;
  double est_bytes;
  double t_submit;
  double deadline;  // wall clock; 0 for none
};

}  // namespace

class Server {
 public:
  Server(int device, int nslots, double budget) : device_(device), budget_(budget) {
    OSH_CHECK(nslots >= 1 && nslots <= 16, "1 to 16 slots");
    OSH_CHECK(budget > 0, "the HBM budget must be positive");
    for (int s = 0; s < nslots; ++s) threads_.emplace_back([this, nslots] { work(nslots); });
  }

  ~Server() {
    {
      std::lock_guard<std::mutex> g(mu_);
      stop_ = true;
    }
    cv_work_.notify_all();
    for (auto& t : threads_) t.join();
  }

  int submit(int dim, int nx, int ny, int nz, int ncomps, double const* metric,
      AdaptOptions const& opts, bool want_digest) {
    OSH_CHECK(dim == 2 || dim == 3, "dimension must be 2 or 3");
    OSH_CHECK(nx > 0 && ny > 0 && (dim == 2 || nz > 0), "cell counts must be positive");
    std::int64_t ncells = std::int64_t(nx) * ny * (dim == 3 ? nz : 1);
    OSH_CHECK(ncells * (dim == 3 ? 6 : 2) < (std::int64_t(1) << 30), "box too large");
    OSH_CHECK(ncomps == 1 || ncomps == (dim == 3 ? 6 : 3), "metric must be isotropic or full");
    Request r;
    r.dim = dim;
    r.nx = nx;
    r.ny = ny;
    r.nz = dim == 3 ? nz : 0;
    r.ncomps = ncomps;
    for (int i = 0; i < ncomps; ++i) r.metric[i] = metric[i];
    r.opts = opts;
    r.want_digest = want_digest;
    double nin = double(ncells) * (dim == 3 ? 6 : 2);
    double nexp = sqrt_det_of_constant_metric(dim, ncomps, metric) / typical_unit_simplex_size(dim);
    r.est_bytes = kBytesPerElement * (nin > nexp ? nin : nexp);
    r.t_submit = wall();
    std::lock_guard<std::mutex> g(mu_);
    r.deadline = timeout_ > 0 ? r.t_submit + timeout_ : 0.0;
    r.ticket = next_ticket_++;
    results_[r.ticket].done = false;
    queue_.push_back(r);
    cv_work_.notify_all();
    return r.ticket;
  }

  // requests submitted from now on fail (status 1) once `seconds` have
  // passed since their submission: before they start, or at the end of the
  // pass that crosses the deadline (a pass itself is not interrupted); 0
  // turns the limit off
  void set_timeout(double seconds) {
    OSH_CHECK(seconds >= 0, "the timeout must not be negative");
    std::lock_guard<std::mutex> g(mu_);
    timeout_ = seconds;
  }

  void wait(int ticket, osh_serve_result* out) {
    std::unique_lock<std::mutex> g(mu_);
    auto it = results_.find(ticket);
    OSH_CHECK(it != results_.end(), "unknown ticket");
    cv_done_.wait(g, [&] { return it->second.done; });
    *out = it->second.r;
    results_.erase(it);
  }

 private:
  struct Outcome {
    bool done = false;
    osh_serve_result r;
  };

  void work(int nslots) {
    std::unique_ptr<Ctx> ctx;
    std::string ctx_error;
    try {
      OSH_HIP(hipSetDevice(device_));
      ctx = std::make_unique<Ctx>();
      // keep each slot's cache of freed blocks within its share of the budget
      ctx->pool->set_cap(std::size_t(budget_ / (2.0 * nslots)));
    } catch (std::exception const& e) {
      ctx_error = e.what();
    }
    for (;;) {
      Request r;
      {
        std::unique_lock<std::mutex> g(mu_);
        cv_work_.wait(g, [&] {
          if (stop_ && queue_.empty()) return true;
          if (queue_.empty()) return false;
          return running_ == 0 || in_use_ + queue_.front().est_bytes <= budget_;
        });
        if (queue_.empty()) return;
        r = queue_.front();
        queue_.pop_front();
        in_use_ += r.est_bytes;
        ++running_;
      }
      osh_serve_result res;
      std::memset(&res, 0, sizeof(res));
      res.ticket = r.ticket;
      double t_start = wall();
      try {
        OSH_CHECK(ctx != nullptr, "this slot has no device context: " + ctx_error);
        OSH_CHECK(r.deadline == 0 || t_start <= r.deadline,
            "request timed out before it started");
        Ctx& c = *ctx;
        DMesh m;
        build_box(c, m, r.dim, 1.0, 1.0, r.dim == 3 ? 1.0 : 0.0, r.nx, r.ny, r.nz, true);
        LO nv = m.nents[0];
        std::vector<Real> met(std::size_t(nv) * r.ncomps);
        for (LO v = 0; v < nv; ++v)
          for (int k = 0; k < r.ncomps; ++k) met[std::size_t(v) * r.ncomps + k] = r.metric[k];
        m.metric = to_device(c, met);
        m.metric_ncomps = r.ncomps;
        res.nelems_in = m.nents[r.dim];
        std::vector<PassRecord> log;
        PassHook hook;
        if (r.deadline > 0)
          hook = [&](DMesh&) {
            OSH_CHECK(wall() <= r.deadline, "request timed out after " +
                                                std::to_string(log.size()) + " passes");
          };
        adapt(c, m, r.opts, &log, &hook);
        c.sync();
        res.nelems_out = m.nents[r.dim];
        res.npasses = int(log.size());
        for (auto const& p : log) {
          res.pass_elems += p.nelems_in;
          res.pass_seconds += p.seconds;
        }
        if (r.want_digest) res.digest = mesh_digest(c, m);
      } catch (std::exception const& e) {
        res.status = 1;
        std::strncpy(res.error, e.what(), sizeof(res.error) - 1);
      }
      double t_done = wall();
      res.queue_seconds = t_start - r.t_submit;
      res.run_seconds = t_done - t_start;
      res.latency_seconds = t_done - r.t_submit;
      {
        std::lock_guard<std::mutex> g(mu_);
        in_use_ -= r.est_bytes;
        --running_;
        auto& o = results_[r.ticket];
        o.r = res;
        o.done = true;
      }
      cv_done_.notify_all();
      cv_work_.notify_all();
    }
  }

  int device_;
  double budget_;
  std::mutex mu_;
  std::condition_variable cv_work_, cv_done_;
  std::deque<Request> queue_;
  std::map<int, Outcome> results_;
  double in_use_ = 0;
  int running_ = 0;
  bool stop_ = false;
  int next_ticket_ = 0;
  double timeout_ = 0;
  std::vector<std::thread> threads_;
};

}  // namespace osh

struct osh_server {
  osh::Server s;
  osh_server(int device, int nslots, double budget) : s(device, nslots, budget) {}
};

extern "C" {

int osh_server_create(int device, int nslots, double hbm_budget_bytes, osh_server** out) {
  return osh::guarded([&] {
    OSH_CHECK(out != nullptr, "null output pointer");
    *out = new osh_server(device, nslots, hbm_budget_bytes);
  });
}

int osh_server_destroy(osh_server* server) {
  return osh::guarded([&] { delete server; });
}

int osh_server_submit(osh_server* server, int dim, int nx, int ny, int nz, int metric_ncomps,
    const double* metric, const osh_adapt_opts* opts, int want_digest, int* ticket) {
  return osh::guarded([&] {
    OSH_CHECK(server && metric && opts && ticket, "null argument");
    osh::AdaptOptions o;
    osh::to_options(opts, o);
    *ticket = server->s.submit(dim, nx, ny, nz, metric_ncomps, metric, o, want_digest != 0);
  });
}

int osh_server_set_timeout(osh_server* server, double seconds) {
  return osh::guarded([&] {
    OSH_CHECK(server != nullptr, "null argument");
    server->s.set_timeout(seconds);
  });
}

int osh_server_wait(osh_server* server, int ticket, osh_serve_result* out) {
  return osh::guarded([&] {
    OSH_CHECK(server && out, "null argument");
    server->s.wait(ticket, out);
  });
}

}  // extern "C"
