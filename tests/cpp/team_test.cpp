// Host thread team (linearalgebra_amd/csrc/team.hpp): a member whose work
// throws aborts the run -- members waiting on the team's barriers leave
// them -- and Team::run rethrows that exception once every member is out;
// the team is usable afterwards. Exit code = number of failed checks.
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../linearalgebra_amd/csrc/team.hpp"

using laamd::team::Barrier;
using laamd::team::Team;

static int failures = 0;
static void expect(bool ok, const char *name) {
  std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
  if (!ok) ++failures;
}

// every member passes `rounds` barriers; member `thrower` throws after
// `after` of them (thrower < 0: nobody throws). Returns the message of the
// exception run() rethrew, or "" when it returned normally.
static std::string run_case(Team &T, int rounds, int thrower, int after, std::vector<int> &count) {
  Barrier bar(T);
  count.assign(T.size(), 0);
  try {
    T.run([&](int t, int) {
      for (int r = 0; r < rounds; ++r) {
        if (t == thrower && r == after) throw std::runtime_error("member failed");
        bar.wait();
        ++count[t];
      }
    });
  } catch (const std::runtime_error &e) {
    return e.what();
  }
  return "";
}

int main() {
  std::vector<int> count;
  {
    Team T(8);
    expect(run_case(T, 50, -1, 0, count) == "", "8 members, no failure");
    bool all = true;
    for (int c : count) all = all && c == 50;
    expect(all, "every member passed every barrier");
    expect(run_case(T, 50, 3, 7, count) == "member failed", "worker throws: run rethrows");
    expect(count[3] == 7, "the thrower stopped where it threw");
    expect(run_case(T, 50, 0, 11, count) == "member failed", "caller (member 0) throws");
    expect(run_case(T, 20, -1, 0, count) == "", "team usable after an aborted run");
    all = true;
    for (int c : count) all = all && c == 20;
    expect(all, "the next run completes every barrier");
  }
  {
    Team T(1);
    expect(run_case(T, 5, 0, 2, count) == "member failed", "single member throws");
    expect(run_case(T, 5, -1, 0, count) == "", "single member, no failure");
  }
  return failures;
}
