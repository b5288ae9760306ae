// CPU cases for the library's host re-validation (path_ok in
// metal-route_amd/csrc/hostcheck.h), the fail-closed check every device path
// passes before it is returned (mr-metal lib.rs:902-974, negotiated.rs:1284-1331).
// Exit status 0 when every case holds; each failure prints its name.
#include <cstdio>

#include "../metal-route_amd/csrc/hostcheck.h"

using namespace mrf::host;

static int fails = 0;

static void expect(bool got, bool want, const char* name) {
  if (got != want) {
    printf("FAIL %s: got %d want %d\n", name, (int)got, (int)want);
    fails++;
  }
}

int main() {
  const uint32_t W = 40, H = 40, L = 2;
  std::vector<uint32_t> cost(W * H * L, 1), xe(W, 16), ye(H, 16), ve(L, 160);
  std::vector<uint8_t> va(L, 1);
  HostGrid g;
  g.d = mrf_dims{W, H, L};
  g.cost = cost.data();
  g.xe = xe.data();
  g.ye = ye.data();
  g.ve = ve.data();
  g.va = va.data();
  NetPlan p;
  const mrf_window all{0, 0, W - 1, H - 1};
  mrf_net n{};
  auto check = [&](const std::vector<uint32_t>& path, uint32_t c, const mrf_window& w = mrf_window{0, 0, 39, 39}) {
    n.src = path.front();
    n.dst = path.back();
    return path_ok(g, n, p, w, -1, path.data(), path.size(), c, false);
  };
  // a snake over the whole top layer: 1600 cells, every one once
  std::vector<uint32_t> snake;
  for (uint32_t y = 0; y < H; ++y)
    for (uint32_t k = 0; k < W; ++k) snake.push_back(y * W + (y % 2 ? W - 1 - k : k));
  const uint32_t sc = (uint32_t)(snake.size() - 1) * 16;
  expect(check(snake, sc), true, "long simple path");
  expect(check(snake, sc + 1), false, "search cost differs");
  // revisits with a matching cost sum
  expect(check({0, 1, 2, 1, 2, 3}, 5 * 16), false, "short revisit");
  std::vector<uint32_t> back(snake.begin(), snake.begin() + 200);
  back.push_back(back.back() - W);  // up into the row already walked
  expect(check(back, (uint32_t)(back.size() - 1) * 16), false, "revisit 200 cells later");
  // adjacency, vias and the window
  expect(check({0, 1, 2, 3}, 48), true, "straight");
  expect(check({0, 2, 3}, 32), false, "jump of two cells");
  expect(check({0, W * H, W * H + 1}, 160 + 16), true, "via then planar");
  expect(check({0, W * H + 1}, 160), false, "diagonal via");
  expect(check({0, 1, 2, 3}, 48, mrf_window{0, 0, 2, 39}), false, "leaves the window");
  // an obstacle is passable only as one of the net's pads, at weight 1
  cost[2] = mrf::kMax;
  expect(check({0, 1, 2, 3}, 48), false, "obstacle");
  p.pads = {2};
  expect(check({0, 1, 2, 3}, 48), true, "own pad");
  p.pads.clear();
  cost[2] = 1;
  // the far corner of the second layer: largest index on the board
  const uint32_t last = W * H * L - 1;
  expect(check({last - 1, last}, 16), true, "last cell");
  expect(check({last, last + 1}, 16), false, "off the board");
  (void)all;
  printf("%s\n", fails ? "hostcheck cases failed" : "hostcheck cases ok");
  return fails ? 1 : 0;
}
