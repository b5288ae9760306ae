// The SimpleRouteJson producer (metal-route_amd/csrc/srj.cpp, host code) as a
// program, so that it runs under the host sanitizers.  For each file named on
// the command line: rasterise it in the three modes and check the outputs'
// shapes (line arrays sorted, nets' cells in range, pads ascending, via pads a
// subset of pads); then feed the parser about 300 evenly spaced prefixes of
// the document and about 100 copies with one byte replaced, each of which must
// either parse or fail with an error message, never crash.  Exit status 0 when
// every case holds.
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iterator>
#include <string>
#include <vector>

#include "../include/mrfield.h"

static int fails = 0;
static void expect(bool ok, const std::string& what) {
  if (!ok) {
    printf("FAIL %s\n", what.c_str());
    fails++;
  }
}

static void check(const mrf_srj_problem* p, const std::string& tag) {
  const size_t n = (size_t)p->dims.w * p->dims.h * p->dims.layers;
  for (uint32_t i = 1; i < p->dims.w; ++i) expect(p->x_lines[i - 1] < p->x_lines[i], tag + " x lines sorted");
  for (uint32_t i = 1; i < p->dims.h; ++i) expect(p->y_lines[i - 1] < p->y_lines[i], tag + " y lines sorted");
  for (size_t c = 0; c < n; ++c) expect(p->cost[c] == 1 || p->cost[c] == 0xFFFFFFFFu, tag + " cost values");
  for (size_t k = 0; k < p->n_nets; ++k) {
    const mrf_srj_net& t = p->nets[k];
    expect(t.src < n && t.dst < n, tag + " endpoints in range");
    for (size_t i = 0; i < t.n_pads; ++i) expect(t.pads[i] < n && (i == 0 || t.pads[i - 1] < t.pads[i]), tag + " pads");
    for (size_t i = 0; i < t.n_via_pads; ++i) {
      bool in = false;
      for (size_t j = 0; j < t.n_pads; ++j) in = in || t.pads[j] == t.via_pads[i];
      expect(in, tag + " via pads within pads");
    }
  }
  for (size_t i = 0; i < p->n_pins; ++i) expect(p->pin_cells[i] < n, tag + " pins in range");
}

int main(int argc, char** argv) {
  for (int a = 1; a < argc; ++a) {
    std::ifstream f(argv[a], std::ios::binary);
    const std::string doc((std::istreambuf_iterator<char>(f)), std::istreambuf_iterator<char>());
    expect(!doc.empty(), std::string(argv[a]) + " readable");
    for (int mode = 0; mode < 3; ++mode) {
      mrf_srj_params prm{};
      prm.mode = mode;
      prm.clearance_cells = 1;
      prm.min_clearance_mm = 0.15;
      prm.via_pad_mm = 0.45;
      mrf_srj_problem* p = nullptr;
      mrf_error err{};
      const int rc = mrf_srj_rasterize(doc.data(), doc.size(), &prm, &p, &err);
      if (mode == MRF_SRJ_PHYSICAL && rc != MRF_OK) continue;  // typed rules absent
      expect(rc == MRF_OK && p, std::string(argv[a]) + " rasterises: " + err.message);
      if (p) check(p, argv[a]);
      mrf_srj_free(p);
    }
    // prefixes, and copies with one byte replaced
    const size_t step = doc.size() / 300 + 1;
    for (size_t cut = 0; cut < doc.size(); cut += step) {
      mrf_srj_problem* p = nullptr;
      mrf_error err{};
      if (mrf_srj_rasterize(doc.data(), cut, nullptr, &p, &err) == MRF_OK) {
        check(p, "prefix");
        mrf_srj_free(p);
      } else {
        expect(err.message[0] != 0, "prefix error has a message");
      }
    }
    const char subst[] = {'"', '{', '}', '[', ']', ',', ':', '\\', '0', 'e', '-', '\x01'};
    for (size_t i = 0; i < doc.size(); i += step * 3 + 1) {
      std::string bad = doc;
      bad[i] = subst[i % sizeof(subst)];
      mrf_srj_problem* p = nullptr;
      mrf_error err{};
      if (mrf_srj_rasterize(bad.data(), bad.size(), nullptr, &p, &err) == MRF_OK) {
        check(p, "corrupted");
        mrf_srj_free(p);
      }
    }
  }
  printf("%s\n", fails ? "FAILED" : "ok");
  return fails ? 1 : 0;
}
