// Timer tree of named ranges and its listings (see osh_profile.hpp); plain
// host C++, so that its tests build without the device toolchain.
#include "osh_profile.hpp"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>

namespace osh {

double profile_now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
      .count();
}

std::size_t Profile::child_of(std::size_t parent, char const* name) {
  std::size_t k = (parent == kNone) ? (frames.empty() ? kNone : 0) : frames[parent].first_child;
  for (; k != kNone; k = frames[k].next_sibling)
    if (frames[k].parent == parent && frames[k].name == name) return k;
  std::size_t id = frames.size();
  Frame f;
  f.parent = parent;
  f.name = name;
  frames.push_back(f);
  if (parent == kNone) {
    if (last_root != kNone) frames[last_root].next_sibling = id;
    last_root = id;
  } else {
    Frame& p = frames[parent];
    if (p.last_child == kNone)
      p.first_child = id;
    else
      frames[p.last_child].next_sibling = id;
    p.last_child = id;
  }
  return id;
}

void Profile::start(char const* name, double now) {
  current = child_of(current, name);
  frames[current].calls += 1;
  frames[current].started = now;
}

void Profile::stop(double now) {
  if (current == kNone) return;
  Frame& f = frames[current];
  f.seconds += now - f.started;
  current = f.parent;
}

void Profile::clear() {
  frames.clear();
  current = kNone;
  last_root = kNone;
}

double Profile::root_seconds() const {
  double t = 0.0;
  for (std::size_t k = frames.empty() ? kNone : 0; k != kNone; k = frames[k].next_sibling)
    t += frames[k].seconds;
  return t;
}

Profile Profile::inverted() const {
  Profile inv;
  // breadth first from the roots, so that the bottom-up roots come in the
  // order in which their names first appear level by level
  std::deque<std::size_t> q;
  for (std::size_t k = frames.empty() ? kNone : 0; k != kNone; k = frames[k].next_sibling)
    q.push_back(k);
  while (!q.empty()) {
    std::size_t node = q.front();
    q.pop_front();
    double self = frames[node].seconds;
    for (std::size_t ch = frames[node].first_child; ch != kNone; ch = frames[ch].next_sibling) {
      self -= frames[ch].seconds;
      q.push_back(ch);
    }
    self = std::max(self, 0.0);
    std::size_t at = kNone;
    for (std::size_t up = node; up != kNone; up = frames[up].parent) {
      at = inv.child_of(at, frames[up].name.c_str());
      inv.frames[at].seconds += self;
      inv.frames[at].calls += frames[node].calls;
    }
  }
  return inv;
}

namespace {

void list(Profile const& p, std::size_t parent, int depth, double chop, bool percent,
    double total, std::string& out) {
  std::vector<std::size_t> kids;
  std::size_t k = (parent == Profile::kNone) ? (p.frames.empty() ? Profile::kNone : 0)
                                             : p.frames[parent].first_child;
  for (; k != Profile::kNone; k = p.frames[k].next_sibling) kids.push_back(k);
  std::stable_sort(kids.begin(), kids.end(),
      [&](std::size_t a, std::size_t b) { return p.frames[a].seconds > p.frames[b].seconds; });
  for (std::size_t ch : kids) {
    Profile::Frame const& f = p.frames[ch];
    double share = total > 0.0 ? 100.0 * f.seconds / total : 100.0;
    if (share >= chop) {
      for (int i = 0; i < depth; ++i) out += "|  ";
      char line[96];
      if (percent)
        std::snprintf(line, sizeof line, " %.2f%% %lld\n", share, f.calls);
      else
        std::snprintf(line, sizeof line, " %.6f %lld\n", f.seconds, f.calls);
      out += f.name;
      out += line;
    }
    list(p, ch, depth + 1, chop, percent, total, out);
  }
}

}  // namespace

std::string Profile::report(double chop, bool percent, double total) const {
  std::string out = "\nTOP-DOWN:\n=========\n";
  list(*this, kNone, 0, chop, percent, total, out);
  out += "\nBOTTOM-UP:\n==========\n";
  Profile inv = inverted();
  list(inv, kNone, 0, chop, percent, total, out);
  return out;
}

void Profile::from_environment() {
  origin = profile_now();
  char const* t = std::getenv("OSH_TIME");
  if (!t || !*t || std::strcmp(t, "0") == 0) return;
  int f = std::atoi(t);
  flags = (f > 1) ? (f & (kProfileTree | kProfileMarkers | kProfileSync))
                  : (kProfileTree | kProfileSync);
  print_at_close = (flags & kProfileTree) != 0;
  char const* p = std::getenv("OSH_TIME_PERCENT");
  percent = p && *p && std::strcmp(p, "0") != 0;
  char const* ch = std::getenv("OSH_TIME_CHOP");
  if (ch && *ch) chop = std::atof(ch);
}

}  // namespace osh
