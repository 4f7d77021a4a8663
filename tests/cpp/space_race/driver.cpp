// Several threads set up and read the device state of one shared space, as concurrent
// assemblies of different operators (Laplace, Helmholtz, Maxwell: far-field tables of several
// degrees, with and without tangents) on handles that share the space do.  Every table must be
// complete when it is handed out and keep its address for the life of the space.
#include <atomic>
#include <cstdio>
#include <map>
#include <mutex>
#include <random>
#include <thread>
#include <vector>

#include "bmi_internal.h"

int main(int argc, char** argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 20;
  std::atomic<int> bad{0};
  for (int round = 0; round < rounds; ++round) {
    bmi_geometry geo;
    geo.patches.resize(6);
    for (int p = 0; p < 6; ++p) {
      geo.patches[p].pdx = geo.patches[p].pdy = 5;
      geo.patches[p].cp.assign(100, p + 0.5);
    }
    auto* s = new bmi_space;
    s->geo = &geo;
    s->M = 2, s->n = 4, s->pd = 5, s->npatch = 6, s->nel = 96;
    s->elem_patch.assign(s->nel, 1);
    s->elem_llc.assign(2 * s->nel, 0.25);
    s->rowptr.assign(10, 3);
    s->col.assign(9, 2), s->val.assign(9, 1.5);
    s->trowptr = s->rowptr, s->tcol = s->col, s->tval = s->val;
    std::mutex seen_mu;
    std::map<std::pair<int, int>, const bmi::FFTable*> seen;
    std::vector<std::thread> th;
    for (int t = 0; t < 8; ++t)
      th.emplace_back([&, t] {
        std::mt19937 g(round * 31 + t);
        hipStream_t st;
        hipStreamCreateWithFlags(&st, 0);
        for (int k = 0; k < 40; ++k) {
          bmi::ensure_device_space(s, st);
          if (s->d_cp.ptr[250] != 2.5 || s->d_val.ptr[8] != 1.5) bad.fetch_add(1);
          const int deg = (int[]){3, 4, 5, 12}[g() % 4];
          const bool tan = g() % 3 == 0;
          const bmi::FFTable* tab = bmi::ensure_ffield(s, deg, st, tan);
          const int n = deg + 1;
          if (tab->deg != deg || (tan && tab->fs != bmi::FF_STRIDE_TAN)) bad.fetch_add(1);
          const int64_t total = static_cast<int64_t>(s->nel) * n * n * tab->fs;
          for (int64_t i = 0; i < total; i += 97)
            if (tab->buf.ptr[i] != n * 100000.0 + tab->fs * 1000.0 + static_cast<double>(i % 997)) bad.fetch_add(1);
          std::lock_guard<std::mutex> lk(seen_mu);
          auto key = std::make_pair(deg, tab->fs);
          auto it = seen.find(key);
          if (it == seen.end()) seen[key] = tab;
          else if (it->second != tab) bad.fetch_add(1);  // a table moved or was made twice
        }
        hipStreamDestroy(st);
      });
    for (auto& x : th) x.join();
    delete s;
  }
  if (bad.load()) {
    std::printf("bad %d\n", bad.load());
    return 1;
  }
  std::printf("ok\n");
  return 0;
}
