// dcd_writer.hpp -- CHARMM/LAMMPS DCD trajectories, as `dump ID all dcd N file`
// writes them (the layout of cases/Diffusion Coefficient/1base.dcd).
//
// All records are Fortran unformatted, little-endian: each record is
// framed by its byte length before and after.
//
// Header, record 1 (84 bytes):
//   "CORD" followed by 20 int32 control words:
//     [0] frames written       [1] first timestep written   [2] steps between frames
//     [3] last timestep written   [4..8] 0   [9] timestep as float32
//     [10] 1 (a unit cell follows in every frame)   [11..18] 0   [19] 24
// Header, record 2 (164 bytes):
//   int32 2, then two 80-character title lines
// Header, record 3 (4 bytes):
//   int32 number of atoms
//
// Each frame:
//   record of 6 float64 cell values (x length, 0, y length, 0, 0, z length)
//   then three records of natoms float32 values (x, y, z)
//
// Words [0] and [3] are rewritten after every frame, so the file is a
// complete trajectory whenever it is read.
#pragma once

#include <cstdio>
#include <string>

namespace selm_host {

class DcdWriter {
  public:
    DcdWriter() = default;
    DcdWriter(const DcdWriter &) = delete;
    DcdWriter &operator=(const DcdWriter &) = delete;
    ~DcdWriter() { close(); }

    // Create (truncate) `path`; `every` steps between frames; dt the timestep.
    bool open(const std::string &path, int natoms, int every, double dt, const std::string &title);
    // Append one frame at `step`: x is [natoms][3] float64, box the edge lengths.
    bool frame(long long step, const double *x, const double box[3]);
    void close();
    int frames() const { return nframes_; }

  private:
    void rewrite_counts();
    std::FILE *fp_ = nullptr;
    int natoms_ = 0, every_ = 0, nframes_ = 0;
    long long first_ = 0, last_ = 0;
};

}  // namespace selm_host
