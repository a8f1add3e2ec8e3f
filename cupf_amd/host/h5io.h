// h5io.h -- minimal HDF5 access for the drop-in's inputs (Temp.h5) and outputs
// (QoI and field files, 3DgrainMPI/PhaseFieldCpu.cpp:464-613).  libhdf5 is
// loaded at run time, so the library itself has no link-time HDF5 dependency;
// without it, 1-D datasets go to <file>.d/<name>.bin (raw little-endian) and
// reading returns false.
#pragma once
#include <string>
#include <vector>

namespace h5io {

bool available();

class Writer {
public:
    explicit Writer(const std::string& path);
    ~Writer();
    bool ok() const { return ok_; }
    bool write_f32(const std::string& name, const float* data, size_t n);
    bool write_i32(const std::string& name, const int* data, size_t n);

private:
    std::string path_;
    long long file_ = -1;
    bool ok_ = false;
    bool raw_ = false;
};

// Read a whole 1-D float dataset; false if the file or dataset is missing.
bool read_f32(const std::string& path, const std::string& name, std::vector<float>& out);

}  // namespace h5io
