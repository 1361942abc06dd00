// cosmo_main.cc -- the `cosmo_gpu` driver: `cosmo_gpu <config file>` runs the
// simulation the config names, like the reference's `cosmo <config>`
// (cosmo.cc:24-100).  `cosmo_gpu --print-params <config file>` prints the
// grid and physics parameters the config resolves to and touches no GPU.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "cosmo_host.h"

int main(int argc, char** argv) {
  const bool print = argc == 3 && !std::strcmp(argv[1], "--print-params");
  if (argc != 2 && !print) {
    std::cout << "Error: please supply exactly one config filename as an argument.\n";
    return EXIT_FAILURE;
  }
  try {
    cosmo::ConfigParser cfg(argv[argc - 1]);
    if (print) {
      const cg_params p = cosmo::params_from_config(cfg);
      std::printf("nx=%d ny=%d nz=%d stencil_order=%d precision=%d dx=%.17g dt=%.17g ko_coeff=%.17g "
                  "a_adj_amp=%.17g k_damping_amp=%.17g normalize_metric=%d lapse=%d matter=%d device=%d\n",
                  p.nx, p.ny, p.nz, p.stencil_order, p.precision, p.dx, p.dt, p.ko_coeff, p.a_adj_amp,
                  p.k_damping_amp, p.normalize_metric, p.lapse, p.matter, p.device);
      return EXIT_SUCCESS;
    }
    cosmo::CosmoSim* sim = cosmo::make_sim(cfg);
    sim->init();
    sim->setICs();
    sim->run();
    delete sim;
  } catch (int code) {
    std::cerr << "cosmo_gpu: stopped with code " << code << "\n";
    return code == 10 ? 10 : (code == 2 ? 2 : EXIT_FAILURE);
  }
  return EXIT_SUCCESS;
}
