// dcd_writer.cpp -- see dcd_writer.hpp for the layout.
#include "dcd_writer.hpp"

#include <cstdint>
#include <cstring>
#include <ctime>
#include <vector>

namespace selm_host {

namespace {

void put_i32(std::FILE *fp, int32_t v) { std::fwrite(&v, 4, 1, fp); }

}  // namespace

bool DcdWriter::open(const std::string &path, int natoms, int every, double dt, const std::string &title) {
    close();
    fp_ = std::fopen(path.c_str(), "wb");
    if (!fp_) return false;
    natoms_ = natoms;
    every_ = every;
    nframes_ = 0;
    first_ = last_ = 0;
    // record 1
    put_i32(fp_, 84);
    std::fwrite("CORD", 1, 4, fp_);
    int32_t ctl[20] = {0};
    ctl[2] = every;
    float fdt = (float)dt;
    std::memcpy(&ctl[9], &fdt, 4);
    ctl[10] = 1;
    ctl[19] = 24;
    std::fwrite(ctl, 4, 20, fp_);
    put_i32(fp_, 84);
    // record 2: two 80-character titles
    put_i32(fp_, 164);
    put_i32(fp_, 2);
    char t1[80], t2[81];
    std::memset(t1, 0, sizeof t1);
    std::strncpy(t1, title.c_str(), 79);
    std::fwrite(t1, 1, 80, fp_);
    std::time_t now = std::time(nullptr);
    char when[64];
    std::strftime(when, sizeof when, "REMARKS Created %d %B,%Y at %H:%M", std::localtime(&now));
    std::snprintf(t2, sizeof t2, "%-80s", when);
    std::fwrite(t2, 1, 80, fp_);
    put_i32(fp_, 164);
    // record 3
    put_i32(fp_, 4);
    put_i32(fp_, natoms);
    put_i32(fp_, 4);
    std::fflush(fp_);
    return true;
}

bool DcdWriter::frame(long long step, const double *x, const double box[3]) {
    if (!fp_) return false;
    const double cell[6] = {box[0], 0.0, box[1], 0.0, 0.0, box[2]};
    put_i32(fp_, 48);
    std::fwrite(cell, 8, 6, fp_);
    put_i32(fp_, 48);
    std::vector<float> c(natoms_);
    for (int a = 0; a < 3; a++) {
        for (int i = 0; i < natoms_; i++) c[i] = (float)x[3 * (size_t)i + a];
        put_i32(fp_, 4 * natoms_);
        std::fwrite(c.data(), 4, natoms_, fp_);
        put_i32(fp_, 4 * natoms_);
    }
    if (nframes_ == 0) first_ = step;
    last_ = step;
    nframes_++;
    rewrite_counts();
    return std::ferror(fp_) == 0;
}

void DcdWriter::rewrite_counts() {
    const long pos = std::ftell(fp_);
    const int32_t n = nframes_, f = (int32_t)first_, l = (int32_t)last_;
    std::fseek(fp_, 8, SEEK_SET);
    std::fwrite(&n, 4, 1, fp_);
    std::fwrite(&f, 4, 1, fp_);
    std::fseek(fp_, 8 + 12, SEEK_SET);
    std::fwrite(&l, 4, 1, fp_);
    std::fseek(fp_, pos, SEEK_SET);
    std::fflush(fp_);
}

void DcdWriter::close() {
    if (fp_) std::fclose(fp_);
    fp_ = nullptr;
}

}  // namespace selm_host
