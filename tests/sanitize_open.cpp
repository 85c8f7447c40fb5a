/*
 * sanitize_open.cpp -- drives libneurogpu's file boundary (ngpu_open: the
 * BasicConstSegP / Sim / AllParams / Stim_meta / Stim_raw / times parsers,
 * the tree checks and the level-schedule builder of ngpu_set_tree) on the
 * data directories given as arguments, for an AddressSanitizer /
 * UndefinedBehaviorSanitizer build of the host code.  Prints one line per
 * directory: the return code and the message.  Test infrastructure only.
 */
#include <cstdio>

#include "../include/neurogpu.h"

int main(int argc, char **argv)
{
    for (int i = 1; i < argc; i++) {
        ngpu_ctx *c = nullptr;
        const int rc = ngpu_open(argv[i], "", -1, 0, &c);
        std::printf("%s\t%d\t%s\n", argv[i], rc, ngpu_last_error(nullptr));
        if (rc == NGPU_OK) ngpu_close(c);
    }
    return 0;
}
