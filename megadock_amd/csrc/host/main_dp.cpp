// megadock-hip-dp: the reference's job-table farm (src/mpidp.cpp:24-129) in one process
// on MI355X: megadock-hip-dp -tb <table> [-lg log] [-rt retries] [-ot k] [-pj threads]
// [other megadock options for every job].
#include <cstdio>

#include "../../../include/mdk_host.h"

int main(int argc, char **argv) {
    std::printf(" MEGADOCK-compatible job farm on MI355X (libmdk)\n");
    if (mdkh_run_table(argc, (const char **)argv)) {
        std::fprintf(stderr, "%s\n", mdkh_last_error());
        return 1;
    }
    return 0;
}
