/*
 * mcxcl.c - command-line front end (reference src/mcxcl.c:23-41):
 * initialise a Config, parse the command line (which reads the input file or
 * a built-in benchmark), run the simulation, report and save, clean up.
 */
#include <stdio.h>
#include <stdlib.h>

#include "mcx_host.h"
#include "mcx_utils.h"

int main(int argc, char* argv[]) {
    Config cfg;
    mcx_initcfg(&cfg);
    mcx_parsecmd(argc, argv, &cfg);

    if (cfg.issave2pt && !cfg.exportfield) {
        size_t len = (size_t)cfg.dim.x * cfg.dim.y * cfg.dim.z * cfg.maxgate;
        cfg.exportfield = (float*)calloc(len * 2, sizeof(float));
    }

    mcx_run_simulation(&cfg, NULL, NULL);
    printf("simulated %zu photons in %u ms (kernel %.2f ms): %.2f photon/ms\n", cfg.nphoton, cfg.runtime,
           cfg.kernelms, cfg.nphoton / (cfg.kernelms > 0 ? cfg.kernelms : 1.f));
    printf("total simulated energy: %.2f\tabsorbed: %5.5f%%\n", cfg.energytot,
           cfg.energytot > 0 ? (cfg.energytot - cfg.energyesc) / cfg.energytot * 100.0 : 0.0);

    if (cfg.issavedet) {
        printf("detected %u photons\n", cfg.detectedcount);
    }

    if (cfg.issave2pt && cfg.exportfield) {
        mcx_savedata(cfg.exportfield, (size_t)cfg.dim.x * cfg.dim.y * cfg.dim.z * cfg.maxgate, &cfg);
    }

    if (cfg.issavedet && cfg.exportdetected) {
        mcx_savedetphoton(cfg.exportdetected, cfg.seeddata, (int)cfg.detectedcount, 0, &cfg);
    }

    mcx_clearcfg(&cfg);
    return 0;
}
