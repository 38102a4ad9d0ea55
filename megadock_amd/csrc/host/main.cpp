// megadock-hip: the reference's command line (src/main.cpp:38-140) on MI355X.
#include <cstdio>

#include "../../../include/mdk_host.h"

int main(int argc, char **argv) {
    std::printf(" MEGADOCK-compatible FFT docking on MI355X (libmdk)\n");
    if (mdkh_run(argc, (const char **)argv)) {
        std::fprintf(stderr, "%s\n", mdkh_last_error());
        return 1;
    }
    return 0;
}
