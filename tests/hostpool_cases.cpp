// CPU cases for the library's shared host workers (metal-route_amd/csrc/
// hostpool.h): every index of a job is processed exactly once, jobs run back
// to back and from two submitting threads at once, a chunk's exception reaches
// the caller, a call made inside a job runs inline, and a forked child starts
// workers of its own.  Exit status 0 when
// every case holds.
#include <atomic>
#include <cstdio>
#include <stdexcept>
#include <thread>
#include <vector>

#include <sys/wait.h>
#include <unistd.h>

#include "../metal-route_amd/csrc/hostpool.h"

using mrf::host::parallel_for;

static int fails = 0;

static void expect(bool ok, const char* name) {
  if (!ok) {
    printf("FAIL %s\n", name);
    fails++;
  }
}

static bool once_each(size_t n, size_t chunk) {
  std::vector<std::atomic<int>> hits(n);
  for (auto& h : hits) h.store(0);
  parallel_for(n, chunk, [&](size_t a, size_t b) {
    for (size_t i = a; i < b; ++i) hits[i].fetch_add(1);
  });
  for (auto& h : hits)
    if (h.load() != 1) return false;
  return true;
}

int main() {
  bool ok = true;
  for (size_t n : {0, 1, 7, 31, 32, 33, 1000, 100000})
    for (size_t chunk : {1, 16, 1 << 16}) ok = ok && once_each(n, chunk);
  expect(ok, "every index once");
  ok = true;
  for (int rep = 0; rep < 2000; ++rep) ok = ok && once_each(257, 16);
  expect(ok, "2000 jobs back to back");
  bool threw = false;
  try {
    parallel_for(1000, 10, [&](size_t a, size_t) {
      if (a >= 500) throw std::runtime_error("chunk failed");
    });
  } catch (const std::runtime_error&) {
    threw = true;
  }
  expect(threw, "exception reaches the caller");
  expect(once_each(1000, 10), "pool usable after an exception");
  std::atomic<int> inner{0};
  parallel_for(64, 4, [&](size_t a, size_t b) {
    parallel_for(b - a, 1, [&](size_t x, size_t y) { inner.fetch_add((int)(y - x)); });
  });
  expect(inner.load() == 64, "nested call runs inline");
  std::atomic<bool> t_ok{true};
  std::thread other([&] {
    for (int rep = 0; rep < 200; ++rep)
      if (!once_each(5000, 64)) t_ok = false;
  });
  bool m_ok = true;
  for (int rep = 0; rep < 200; ++rep) m_ok = m_ok && once_each(3000, 32);
  other.join();
  expect(m_ok && t_ok.load(), "two submitting threads");
#if !defined(__SANITIZE_THREAD__)  // the thread sanitizer stops threads started after a fork
  // a forked child (the parent's workers exist, the child's do not) runs jobs
  const pid_t child = fork();
  if (child == 0) _exit(once_each(10000, 16) && once_each(5000, 1) ? 0 : 1);
  int status = 0;
  waitpid(child, &status, 0);
  expect(WIFEXITED(status) && WEXITSTATUS(status) == 0, "forked child");
  expect(once_each(10000, 16), "parent after a fork");
#endif
  printf("%s\n", fails ? "hostpool cases failed" : "hostpool cases ok");
  return fails ? 1 : 0;
}
