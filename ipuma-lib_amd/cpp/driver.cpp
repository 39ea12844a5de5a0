// driver.cpp -- one process-global engine (src/driver.cpp:5-17).
#include "driver.hpp"

#include <memory>
#include <mutex>

namespace {
std::unique_ptr<ipu::batchaffine::SWAlgorithm> g_driver;
std::mutex g_mutex;
}  // namespace

ipu::batchaffine::SWAlgorithm *getDriver() { return g_driver.get(); }

void init_single_ipu(ipu::SWConfig config, ipu::IPUAlgoConfig algoconfig) {
  std::lock_guard<std::mutex> lock(g_mutex);
  if (!g_driver) g_driver.reset(new ipu::batchaffine::SWAlgorithm(config, algoconfig));
}
