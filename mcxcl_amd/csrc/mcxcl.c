/*
 * mcxcl.c - command-line front end (reference src/mcxcl.c:23-41):
 * initialise a Config, parse the command line (which reads the input file or
 * a built-in benchmark), run the simulation (which writes the output files of
 * a stand-alone caller), report, clean up.
 */
#include <stdio.h>
#include <stdlib.h>

#include "mcx_host.h"
#include "mcx_utils.h"

int main(int argc, char* argv[]) {
    Config cfg;
    mcx_initcfg(&cfg);
    mcx_parsecmd(argc, argv, &cfg);

    mcx_run_simulation(&cfg, NULL, NULL);
    printf("simulated %zu photons in %u ms (kernel %.2f ms): %.2f photon/ms\n", cfg.nphoton, cfg.runtime,
           cfg.kernelms, cfg.nphoton / (cfg.kernelms > 0 ? cfg.kernelms : 1.f));
    printf("total simulated energy: %.2f\tabsorbed: %5.5f%%\n", cfg.energytot,
           cfg.energytot > 0 ? (cfg.energytot - cfg.energyesc) / cfg.energytot * 100.0 : 0.0);

    if (cfg.issavedet) {
        printf("detected %u photons\n", cfg.detectedcount);
    }

    mcx_clearcfg(&cfg);
    return 0;
}
