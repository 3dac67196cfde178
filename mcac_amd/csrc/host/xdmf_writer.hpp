// Snapshot files in the reference's output format (src/io/writer.cpp,
// src/io/format.cpp, src/io/aggregat_list.cpp, src/io/sphere_list.cpp,
// src/io/physical_model.cpp): `<prefix>_<zero-padded file number>.{h5,xmf}`,
// padding ceil(log10 N) + 4 digits, one temporal collection of Polyvertex
// grids per file, n_time_per_file grids per file.  Array data go to top-level
// HDF5 datasets Data0, Data1, ... referenced from the XML as `<file>.h5:DataK`
// (what the Xdmf library writes and pymcac's readers resolve).  Written with
// the HDF5 C API; the XML is produced here.
#pragma once
#include <cstdint>
#include <string>
#include <utility>
#include <vector>

namespace mcac {

class SnapshotWriter {
  public:
    struct Field {
        std::string name;
        int kind;  // 0 double, 1 int32, 2 int64
        const void *data;
        size_t n;
    };
    // prefix: output path without suffix ("dir/Spheres"); n: number of
    // monomers (file-name width); physics: the "Physics" information block
    SnapshotWriter(std::string prefix, size_t n, size_t per_file) : prefix_(std::move(prefix)), n_(n), per_file_(per_file ? per_file : 1) {}
    ~SnapshotWriter();
    SnapshotWriter(const SnapshotWriter &) = delete;
    SnapshotWriter &operator=(const SnapshotWriter &) = delete;
    // one grid: time, positions (3 n interleaved), node fields
    void write(const std::string &grid_name, double time, const std::vector<double> &pos3,
               const std::vector<Field> &fields, const std::vector<std::pair<std::string, std::string>> &physics);
    void close();  // writes a partly filled file

  private:
    std::string prefix_;
    size_t n_, per_file_;
    size_t step_ = 0, file_ = 0, data_k_ = 0;
    int64_t h5_ = -1;  // hid_t of the open .h5 file
    std::string xml_grids_, xml_head_, base_;
    int64_t dataset(const void *data, size_t n, int kind);
    void open_file(const std::vector<std::pair<std::string, std::string>> &physics);
    void finish_file();
};

}  // namespace mcac
