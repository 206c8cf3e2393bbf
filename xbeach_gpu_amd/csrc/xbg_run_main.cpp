// xbg_run: run the case of XBG_param.txt (or the file given as argument),
// as the reference's XBeach_GPU executable does (Wave_gpu.cu:1639).
#include <cstdio>

#include "../../include/xbg.h"

int main(int argc, char** argv) {
    const char* pf = argc > 1 ? argv[1] : "XBG_param.txt";
    int rc = xbg_run_case(pf, 1);
    if (rc) {
        fprintf(stderr, "xbg_run: %s\n", xbg_run_error());
        return 1;
    }
    return 0;
}
