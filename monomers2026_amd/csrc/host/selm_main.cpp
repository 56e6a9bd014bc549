// selm_main.cpp -- the reference's main (cuda/main.cu:5-12) for the HIP build:
// run the SELM deck in a directory (Model.SELM + Model.LAMMPS_script).
//
//     selm_hip_driver [deck_dir]
#include <cstdio>
#include <exception>

#include "SELM.hpp"

int main(int argc, char **argv) {
    const char *dir = argc > 1 ? argv[1] : ".";
    try {
        selm_host::Selm selm(dir);
        selm.initial();
        selm.run();
        selm.final();
    } catch (const std::exception &e) {
        std::fprintf(stderr, "selm_hip_driver: %s\n", e.what());
        return 1;
    }
    return 0;
}
