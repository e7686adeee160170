// The timer tree of named ranges on a fixed sequence of starts and stops
// with chosen clock readings, against listings worked out by hand: the
// top-down tree sums each chain's time and calls, siblings largest first;
// the bottom-up tree roots each range name at its self time (time less its
// children's) and calls summed over every chain it appears in.
#include <cstdio>
#include <string>

#include "osh_profile.hpp"

static int failures = 0;

static void expect(std::string const& got, std::string const& want, char const* what) {
  if (got != want) {
    std::printf("%s differs\n--- got:\n%s\n--- want:\n%s\n", what, got.c_str(), want.c_str());
    ++failures;
  }
}

int main() {
  osh::Profile p;
  p.start("adapt", 0.0);
  p.start("satisfy_lengths", 1.0);
  p.start("coarsen_by_size", 1.0);
  p.stop(3.0);
  p.start("coarsen_by_size", 3.0);
  p.start("find_indset", 3.25);
  p.stop(3.5);
  p.stop(4.0);
  p.stop(5.0);
  p.start("satisfy_quality", 5.0);
  p.start("swap_edges", 5.0);
  p.start("find_indset", 6.0);
  p.stop(6.5);
  p.stop(7.0);
  p.stop(8.0);
  p.stop(10.0);
  if (p.current != osh::Profile::kNone || p.frames.size() != 7 || p.root_seconds() != 10.0) {
    std::printf("tree shape: %zu frames, root time %g\n", p.frames.size(), p.root_seconds());
    ++failures;
  }
  expect(p.report(0.0, false, 10.0),
      "\nTOP-DOWN:\n=========\n"
      "adapt 10.000000 1\n"
      "|  satisfy_lengths 4.000000 1\n"
      "|  |  coarsen_by_size 3.000000 2\n"
      "|  |  |  find_indset 0.250000 1\n"
      "|  satisfy_quality 3.000000 1\n"
      "|  |  swap_edges 2.000000 1\n"
      "|  |  |  find_indset 0.500000 1\n"
      "\nBOTTOM-UP:\n==========\n"
      "adapt 3.000000 1\n"
      "coarsen_by_size 2.750000 2\n"
      "|  satisfy_lengths 2.750000 2\n"
      "|  |  adapt 2.750000 2\n"
      "swap_edges 1.500000 1\n"
      "|  satisfy_quality 1.500000 1\n"
      "|  |  adapt 1.500000 1\n"
      "satisfy_lengths 1.000000 1\n"
      "|  adapt 1.000000 1\n"
      "satisfy_quality 1.000000 1\n"
      "|  adapt 1.000000 1\n"
      "find_indset 0.750000 2\n"
      "|  swap_edges 0.500000 1\n"
      "|  |  satisfy_quality 0.500000 1\n"
      "|  |  |  adapt 0.500000 1\n"
      "|  coarsen_by_size 0.250000 1\n"
      "|  |  satisfy_lengths 0.250000 1\n"
      "|  |  |  adapt 0.250000 1\n",
      "seconds listing");
  // percentages of a 20 s run, lines under 5 % left out (their children
  // still listed)
  expect(p.report(5.0, true, 20.0),
      "\nTOP-DOWN:\n=========\n"
      "adapt 50.00% 1\n"
      "|  satisfy_lengths 20.00% 1\n"
      "|  |  coarsen_by_size 15.00% 2\n"
      "|  satisfy_quality 15.00% 1\n"
      "|  |  swap_edges 10.00% 1\n"
      "\nBOTTOM-UP:\n==========\n"
      "adapt 15.00% 1\n"
      "coarsen_by_size 13.75% 2\n"
      "|  satisfy_lengths 13.75% 2\n"
      "|  |  adapt 13.75% 2\n"
      "swap_edges 7.50% 1\n"
      "|  satisfy_quality 7.50% 1\n"
      "|  |  adapt 7.50% 1\n"
      "satisfy_lengths 5.00% 1\n"
      "|  adapt 5.00% 1\n"
      "satisfy_quality 5.00% 1\n"
      "|  adapt 5.00% 1\n",
      "percent listing");
  // a second root, and clear()
  p.start("build_box", 10.0);
  p.stop(10.5);
  if (p.root_seconds() != 10.5 || p.frames[0].next_sibling != 7) {
    std::printf("second root not linked\n");
    ++failures;
  }
  p.clear();
  if (!p.frames.empty() || p.root_seconds() != 0.0) ++failures;
  std::printf(failures ? "FAILED\n" : "ok\n");
  return failures ? 1 : 0;
}
