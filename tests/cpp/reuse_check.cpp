// reuse_check.cpp -- unit check of fx::solver_reuse_ok (host-only header
// faspxx_amd/csrc/fx_reuse.hpp): a solver set up again keeps its fused-path
// workspace only for the very handles (address and id) it was set up with.
// Exit status = failures.
#include <cstdio>

#include "fx_reuse.hpp"

struct Mat { uint64_t gen; };
struct Pc { uint64_t gen; };

static int g_fail = 0;
#define CHECK(c)                                              \
  do {                                                        \
    if (!(c)) {                                               \
      std::printf("FAIL %s:%d: %s\n", __FILE__, __LINE__, #c); \
      ++g_fail;                                               \
    }                                                         \
  } while (0)

int main() {
  using fx::solver_reuse_ok;
  Mat a{1}, b{2};
  Pc p{3}, q{4};
  // first setup: nothing to keep
  CHECK(!(solver_reuse_ok<Mat, Pc>(nullptr, 0, nullptr, 0, &a, &p)));
  // the same handles again
  CHECK(solver_reuse_ok(&a, 1, &p, 3, &a, &p));
  // another matrix or preconditioner
  CHECK(!solver_reuse_ok(&a, 1, &p, 3, &b, &p));
  CHECK(!solver_reuse_ok(&a, 1, &p, 3, &a, &q));
  // a matrix freed and made again at the same address carries a new id
  Mat remade{5};
  Mat* same_addr = &a;
  *same_addr = remade;
  CHECK(!solver_reuse_ok(&a, 1, &p, 3, same_addr, &p));
  Pc premade{6};
  p = premade;
  CHECK(!solver_reuse_ok(&a, 5, &p, 3, &a, &p));
  CHECK(solver_reuse_ok(&a, 5, &p, 6, &a, &p));
  std::printf("%s (%d failures)\n", g_fail ? "FAILED" : "PASSED", g_fail);
  return g_fail;
}
