// many clients submitting and waiting concurrently; every result is checked
#include <cstdio>
#include <cstdlib>
#include <random>
#include <thread>
#include <vector>
#include <atomic>
#include "bembel_amd.h"
extern std::atomic<bool> g_teardown;
extern std::atomic<int> g_wrong_thread;
extern std::atomic<bool> g_destroyed;
extern std::atomic<int> g_after_destroy;
void reset_device_owner();
int main(int argc, char** argv) {
  const int rounds = argc > 1 ? atoi(argv[1]) : 20;
  for (int round = 0; round < rounds; ++round) {
    bmi_server_config cfg;
    bmi_server_config_default(&cfg);
    cfg.max_active = 1 + round % 4;
    cfg.max_batch_rhs = round % 3 == 0 ? 1 : 64;
    cfg.hbm_budget_bytes = round % 5 == 0 ? 3e7 : 1e12;  // small budgets force eviction
    cfg.slice_iterations = round % 2 ? 30 : 0;              // GMRES batches overlapping, or one at a time
    bmi_server* srv;
    if (bmi_server_create(&cfg, &srv) != BMI_OK) return 1;
    std::vector<std::thread> clients;
    int bad = 0;
    for (int c = 0; c < 4; ++c)
      clients.emplace_back([&, c] {
        std::mt19937 g(round * 10 + c);
        for (int k = 0; k < 12; ++k) {
          bmi_request r{};
          r.geometry_path = "x.dat";
          r.scale[0] = 1.0 + (g() % 3);  // few distinct operators: sharing, batching, reuse
          r.refinement_level = 2 + g() % 3;
          r.polynomial_degree = 1 + g() % 2;
          bmi_op_desc_default(g() % 2 ? BMI_HELMHOLTZ_SL : BMI_LAPLACE_SL, &r.op);
          r.op.kappa_re = 2.0;
          r.data_fn = r.op.kind == BMI_LAPLACE_SL ? 0 : 1;
          r.nrhs = 1 + g() % 4;
          std::vector<double> p(3 * r.nrhs);
          for (int j = 0; j < r.nrhs; ++j) p[3 * j] = j + 1.0, p[3 * j + 1] = p[3 * j + 2] = 0.0;
          r.data_params = p.data();
          r.data_kappa = 2.0;
          r.tol = 1e-10;
          r.restart = 30;
          r.maxit = 100;
          int64_t id;
          if (bmi_server_submit(srv, &r, &id) != BMI_OK) { __atomic_add_fetch(&bad, 1, __ATOMIC_RELAXED); continue; }
          bmi_result_info info;
          if (bmi_server_wait(srv, id, -1.0, &info) != BMI_OK) { __atomic_add_fetch(&bad, 1, __ATOMIC_RELAXED); continue; }
          if (info.status != BMI_OK && info.status != BMI_ADMISSION_REJECTED) __atomic_add_fetch(&bad, 1, __ATOMIC_RELAXED);
          const int S = r.op.kind == BMI_LAPLACE_SL ? 1 : 2;
          std::vector<double> x(S * info.ndof * r.nrhs + 1);
          bmi_server_fetch(srv, id, x.data());
          if (info.status == BMI_OK)
            for (int64_t i = 0; i < S * info.ndof * r.nrhs; ++i)
              if (x[i] != 0.5 * p[(i % r.nrhs) * 3]) { __atomic_add_fetch(&bad, 1, __ATOMIC_RELAXED); break; }
        }
      });
    for (auto& t : clients) t.join();
    g_teardown = true;
    bmi_server_destroy(srv);
    g_teardown = false;
    // nothing of the server may still run once destroy returned (a thread that outlives it can be
    // inside a device call while the process exits and the runtime is torn down)
    g_destroyed = true;
    std::this_thread::sleep_for(std::chrono::milliseconds(20));
    g_destroyed = false;
    if (g_after_destroy.load()) { printf("round %d: device calls after bmi_server_destroy\n", round); return 4; }
    reset_device_owner();
    if (bad) { printf("round %d: %d bad results\n", round, bad); return 2; }
    if (g_wrong_thread.load()) { printf("round %d: device calls from more than one thread\n", round); return 3; }
  }
  printf("ok\n");
  return 0;
}
