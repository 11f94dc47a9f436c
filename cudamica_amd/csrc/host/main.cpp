/*
 * main.cpp -- the cudaamica command line program: `cudaamica -f <file>`
 * (fraimondo/cudamica src/cudaamica.cu:17-43): parse the parameter file,
 * pick the GPU, partition the samples and run AMICA.
 */
#include <cstdio>

#include "../../../include/amica.h"
#include "../../../include/config.h"
#include "../../../include/device.h"
#include "../../../include/error.h"
#include "../../../include/helpers.h"
#include "../../../include/params.h"

int main(int argc, char* argv[]) {
    programName = argv[0];
    if (!isParam((char*)"-f", argv, argc) || !getParam((char*)"-f", argv, argc)) {
        printf("\nERROR::Script configuration file is mandatory\n\n\n");
        help();
        return ERRORNOFILE;
    }
    char* filename = getParam((char*)"-f", argv, argc);
    C_CHECK_RETURN(getConfig(filename));
    printConfig();
    C_CHECK_RETURN(getDevices());
    C_CHECK_RETURN(selectDevice(currentConfig.devs, currentConfig.devcount));
    C_CHECK_RETURN(recalc_config());
    error r = runamica();
    freeConfig();
    return r == SUCCESS ? 0 : 1;
}
