// Named ranges of one context: a tree of timers keyed by the chain of range
// names that were open (Omega_h::profile::History and ScopedTimer,
// src/Omega_h_profile.hpp:47-232), and roctx ranges that rocprofv3's marker
// trace shows beside the kernels.
//
// Launches are asynchronous, so a range's host time would otherwise charge
// the device work of one stage to whichever later stage happens to wait for
// it.  With kProfileSync each range waits for its context's stream when it
// ends, so a range's time covers the device work that it enqueued; this
// serialises the host and the device and is only for profiling runs.  All
// flags off (the default) costs one branch per range.
#pragma once

#include <cstddef>
#include <string>
#include <vector>

namespace osh {

constexpr int kProfileTree = 1;     // keep the timer tree
constexpr int kProfileMarkers = 2;  // roctx ranges (rocprofv3 --marker-trace)
constexpr int kProfileSync = 4;     // each range waits for the stream at its end

struct Profile {
  static constexpr std::size_t kNone = std::size_t(-1);
  struct Frame {
    std::size_t parent = kNone;
    std::size_t first_child = kNone;
    std::size_t last_child = kNone;
    std::size_t next_sibling = kNone;
    std::string name;
    double started = 0.0;
    double seconds = 0.0;
    long long calls = 0;
  };
  int flags = 0;
  // the options of Omega_h::Library's --osh-time, --osh-time-percent and
  // --osh-time-chop (src/Omega_h_library.cpp:99-131), read from the
  // environment when a context opens (OSH_TIME=1 keeps the tree and waits at
  // range ends, OSH_TIME=<flags> sets the flags, OSH_TIME_PERCENT=1,
  // OSH_TIME_CHOP=<percent>); the listings are printed when the context
  // closes, as ~Library prints them (:191-203)
  bool print_at_close = false;
  bool percent = false;
  double chop = 0.0;
  double origin = 0.0;  // when the context opened
  void from_environment();
  std::vector<Frame> frames;
  std::size_t current = kNone;  // the innermost open range
  std::size_t last_root = kNone;
  // the child of `parent` (kNone: a root) called `name`, made if missing;
  // children keep the order in which they first opened
  std::size_t child_of(std::size_t parent, char const* name);
  void start(char const* name, double now);
  void stop(double now);
  void clear();
  // the bottom-up tree: a root per range name holding that range's self
  // time (its time less its children's, never below zero) and calls, under
  // it the chains of callers that it was reached through (the inversion of
  // Omega_h::profile::invert, src/Omega_h_profile.cpp:125-149)
  Profile inverted() const;
  double root_seconds() const;
  // "TOP-DOWN:" and "BOTTOM-UP:" listings as print_top_down_and_bottom_up
  // prints them (src/Omega_h_profile.cpp:151-214): one line per range,
  // "|  " per level of depth, then name, time and calls, siblings by time,
  // largest first; a line under `chop` percent of `total` is left out (its
  // children are still listed); with `percent` times are percentages of
  // `total`
  std::string report(double chop, bool percent, double total) const;
};

double profile_now();

}  // namespace osh
