#include "xdmf_writer.hpp"

#include <hdf5.h>

#include <cmath>
#include <cstdio>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace mcac {
namespace {

// format.cpp:58-64: "_" + file number padded to ceil(log10 N) + 4 digits
std::string file_suffix(size_t num, size_t n) {
    const int width = static_cast<int>(std::ceil(std::log10(static_cast<float>(n)))) + 4;
    std::ostringstream s;
    s << "_" << std::setfill('0') << std::setw(width) << num;
    return s.str();
}

std::string basename_of(const std::string &path) {
    const size_t p = path.find_last_of('/');
    return p == std::string::npos ? path : path.substr(p + 1);
}

std::string xml_escape(const std::string &s) {
    std::string o;
    for (char c : s) {
        if (c == '&') o += "&amp;";
        else if (c == '<') o += "&lt;";
        else if (c == '>') o += "&gt;";
        else if (c == '"') o += "&quot;";
        else o += c;
    }
    return o;
}

std::string num17(double v) {
    char b[40];
    std::snprintf(b, sizeof b, "%.17g", v);
    return b;
}

const char *type_attrs(int kind) {
    return kind == 0 ? "DataType=\"Float\" Precision=\"8\"" : (kind == 1 ? "DataType=\"Int\" Precision=\"4\""
                                                                          : "DataType=\"Int\" Precision=\"8\"");
}

}  // namespace

SnapshotWriter::~SnapshotWriter() {
    try {
        close();
    } catch (...) {
    }
}

int64_t SnapshotWriter::dataset(const void *data, size_t n, int kind) {
    const std::string name = "Data" + std::to_string(data_k_++);
    hsize_t dims[1] = {static_cast<hsize_t>(n)};
    hid_t space = H5Screate_simple(1, dims, nullptr);
    hid_t type = kind == 0 ? H5T_NATIVE_DOUBLE : (kind == 1 ? H5T_NATIVE_INT32 : H5T_NATIVE_INT64);
    hid_t ds = H5Dcreate2(static_cast<hid_t>(h5_), name.c_str(), type, space, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
    if (ds < 0) {
        H5Sclose(space);
        throw std::runtime_error("HDF5: cannot create " + name);
    }
    herr_t rc = n ? H5Dwrite(ds, type, H5S_ALL, H5S_ALL, H5P_DEFAULT, data) : 0;
    H5Dclose(ds);
    H5Sclose(space);
    if (rc < 0) throw std::runtime_error("HDF5: cannot write " + name);
    return static_cast<int64_t>(data_k_ - 1);
}

void SnapshotWriter::open_file(const std::vector<std::pair<std::string, std::string>> &physics) {
    base_ = prefix_ + file_suffix(file_, n_);
    hid_t f = H5Fcreate((base_ + ".h5").c_str(), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
    if (f < 0) throw std::runtime_error("HDF5: cannot create " + base_ + ".h5");
    h5_ = static_cast<int64_t>(f);
    data_k_ = 0;
    xml_grids_.clear();
    // writer.cpp:78-84: copyright and the physical model's properties in every file
    std::ostringstream h;
    h << "    <Information Name=\"Copyright\" Value=\"Produced by MCAC\"/>\n";
    h << "    <Information Name=\"Physics\" Value=\"Physical properties of the simulation\">\n";
    for (const auto &kv : physics)
        h << "      <Information Name=\"" << xml_escape(kv.first) << "\" Value=\"" << xml_escape(kv.second) << "\"/>\n";
    h << "    </Information>\n";
    xml_head_ = h.str();
}

void SnapshotWriter::finish_file() {
    if (h5_ < 0) return;
    H5Fclose(static_cast<hid_t>(h5_));
    h5_ = -1;
    std::ofstream x(base_ + ".xmf");
    x << "<?xml version=\"1.0\" encoding=\"utf-8\"?>\n";
    x << "<!DOCTYPE Xdmf SYSTEM \"Xdmf.dtd\" []>\n";
    x << "<Xdmf xmlns:xi=\"http://www.w3.org/2001/XInclude\" Version=\"3.0\">\n  <Domain>\n";
    x << xml_head_;
    x << "    <Grid Name=\"Collection\" GridType=\"Collection\" CollectionType=\"Temporal\">\n";
    x << xml_grids_;
    x << "    </Grid>\n  </Domain>\n</Xdmf>\n";
    if (!x) throw std::runtime_error("cannot write " + base_ + ".xmf");
    file_ += 1;
}

void SnapshotWriter::write(const std::string &grid_name, double time, const std::vector<double> &pos3,
                           const std::vector<Field> &fields,
                           const std::vector<std::pair<std::string, std::string>> &physics) {
    if (step_ % per_file_ == 0) open_file(physics);  // writer.cpp:93-96
    const std::string h5name = basename_of(base_) + ".h5";
    const size_t n = pos3.size() / 3;
    std::ostringstream g;
    g << "      <Grid Name=\"" << xml_escape(grid_name) << "\" GridType=\"Uniform\">\n";
    g << "        <Time Value=\"" << num17(time) << "\"/>\n";
    g << "        <Topology TopologyType=\"Polyvertex\" NumberOfElements=\"" << n << "\" NodesPerElement=\"1\"/>\n";
    const int64_t k_pos = dataset(pos3.data(), pos3.size(), 0);
    g << "        <Geometry GeometryType=\"XYZ\">\n          <DataItem " << type_attrs(0) << " Dimensions=\""
      << pos3.size() << "\" Format=\"HDF\">" << h5name << ":Data" << k_pos << "</DataItem>\n        </Geometry>\n";
    // the "Time" attribute (aggregat_list.cpp:30, sphere_list.cpp:36)
    const int64_t k_time = dataset(&time, 1, 0);
    g << "        <Attribute Name=\"Time\" AttributeType=\"None\" Center=\"Grid\">\n          <DataItem "
      << type_attrs(0) << " Dimensions=\"1\" Format=\"HDF\">" << h5name << ":Data" << k_time
      << "</DataItem>\n        </Attribute>\n";
    for (const Field &f : fields) {
        const int64_t k = dataset(f.data, f.n, f.kind);
        g << "        <Attribute Name=\"" << xml_escape(f.name) << "\" AttributeType=\"Scalar\" Center=\"Node\">\n"
          << "          <DataItem " << type_attrs(f.kind) << " Dimensions=\"" << f.n << "\" Format=\"HDF\">" << h5name
          << ":Data" << k << "</DataItem>\n        </Attribute>\n";
    }
    g << "      </Grid>\n";
    xml_grids_ += g.str();
    step_ += 1;
    if (step_ % per_file_ == 0) finish_file();  // writer.cpp:101-106
}

void SnapshotWriter::close() { finish_file(); }  // threaded_io.cpp:42-44: a partly filled file is written

}  // namespace mcac
