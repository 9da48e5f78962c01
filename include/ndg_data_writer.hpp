// ndg_data_writer.hpp -- header-only SEM::Helpers::DataWriter_t without the VTK library.
//
// Mirrors the reference's writer (src/helpers/DataWriter_t.cpp:18-546,
// include/helpers/DataWriter_t.h:14-33): DataWriter_t(path) creates (or empties) "<path>.series"
// and "<dir>/<stem>_element<ext>.series"; write_data and write_complete_data write one file per
// call named "<stem>_t<time, 9 decimals, '.' -> '_'>s<ext>" in the path's directory and append
// {"name", "time"} to the series. Layout as the reference's: the n x n output points of every
// element (z = 0), one VTK_QUAD per output sub-cell with points (i+1, j), (i+1, j+1), (i, j+1),
// (i, j); point data "Pressure", "Velocity" (2 components), and for write_complete_data
// "PressureDerivative", "VelocityDerivative", "PressureAnalyticalError",
// "VelocityAnalyticalError", plus the "_element" file (one quad per element over its output
// corners, cell data N, Index, Status, Rotation, Rank, PressureError, VelocityError, PressureSigma,
// VelocitySigma, Refine, Coarsen, SplitLevel).
//
// ".vtu" writes one UnstructuredGrid file; ".pvtu" writes a PUnstructuredGrid file naming one
// piece, "<file stem>_0.vtu", beside it. Arrays are inline, uncompressed base64 ("binary": a
// UInt64 byte count, then the little-endian values), points Float64. The series list is kept in
// memory and the file rewritten on each call (the reference re-reads it; the content is the same
// for one writer).
#ifndef NDG_DATA_WRITER_HPP
#define NDG_DATA_WRITER_HPP

#include <algorithm>
#include <charconv>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

namespace SEM {
namespace Helpers {

namespace detail {

inline std::string base64(const std::vector<unsigned char> &in) {
    static const char table[] = "ABCDEFGHIJKLMNOPQRSTUVWXYZabcdefghijklmnopqrstuvwxyz0123456789+/";
    std::string out;
    out.reserve((in.size() + 2) / 3 * 4);
    std::size_t i = 0;
    for (; i + 3 <= in.size(); i += 3) {
        const uint32_t v = (uint32_t)in[i] << 16 | (uint32_t)in[i + 1] << 8 | in[i + 2];
        out += table[v >> 18 & 63];
        out += table[v >> 12 & 63];
        out += table[v >> 6 & 63];
        out += table[v & 63];
    }
    if (i + 1 == in.size()) {
        const uint32_t v = (uint32_t)in[i] << 16;
        out += table[v >> 18 & 63];
        out += table[v >> 12 & 63];
        out += "==";
    } else if (i + 2 == in.size()) {
        const uint32_t v = (uint32_t)in[i] << 16 | (uint32_t)in[i + 1] << 8;
        out += table[v >> 18 & 63];
        out += table[v >> 12 & 63];
        out += table[v >> 6 & 63];
        out += '=';
    }
    return out;
}

// Shortest round-trip text of a double, with ".0" on integral values (as JSON writers print).
inline std::string number(double v, bool json) {
    char buf[64];
    const auto r = std::to_chars(buf, buf + sizeof buf, v);
    std::string s(buf, r.ptr);
    if (json && s.find_first_of(".eEn") == std::string::npos) s += ".0";
    return s;
}

template <typename T> struct VtkType;
template <> struct VtkType<double> { static constexpr const char *name = "Float64"; };
template <> struct VtkType<int32_t> { static constexpr const char *name = "Int32"; };
template <> struct VtkType<int64_t> { static constexpr const char *name = "Int64"; };
template <> struct VtkType<uint64_t> { static constexpr const char *name = "UInt64"; };
template <> struct VtkType<uint8_t> { static constexpr const char *name = "UInt8"; };

// One named array of `components` values per tuple.
struct Array {
    std::string name, type;
    int components = 1;
    std::string text;  // encoded payload
};

template <typename T>
Array make_array(const std::string &name, const std::vector<T> &values, int components, bool ascii) {
    Array a{name, VtkType<T>::name, components, {}};
    if (ascii) {
        std::ostringstream s;
        for (std::size_t i = 0; i < values.size(); ++i) {
            if (i) s << ' ';
            if constexpr (std::is_floating_point_v<T>) s << number(values[i], false);
            else s << +values[i];
        }
        a.text = s.str();
    } else {
        const uint64_t nbytes = values.size() * sizeof(T);
        std::vector<unsigned char> raw(8 + nbytes);
        std::memcpy(raw.data(), &nbytes, 8);  // little-endian host (x86-64)
        if (nbytes) std::memcpy(raw.data() + 8, values.data(), nbytes);
        a.text = base64(raw);
    }
    return a;
}

inline std::string attrs(const Array &a) {
    std::string s = "type=\"" + a.type + "\" Name=\"" + a.name + "\"";
    if (a.components != 1) s += " NumberOfComponents=\"" + std::to_string(a.components) + "\"";
    return s;
}

inline void data_array(std::ostream &o, const Array &a, bool ascii) {
    o << "        <DataArray " << attrs(a) << " format=\"" << (ascii ? "ascii" : "binary") << "\">\n          "
      << a.text << "\n        </DataArray>\n";
}

inline std::vector<double> interleave(const std::vector<double> &a, const std::vector<double> &b) {
    std::vector<double> out(2 * a.size());
    for (std::size_t i = 0; i < a.size(); ++i) {
        out[2 * i] = a[i];
        out[2 * i + 1] = b[i];
    }
    return out;
}

}  // namespace detail

class DataWriter_t {
  public:
    explicit DataWriter_t(std::filesystem::path output_filename, bool ascii = false)
        : directory_(output_filename.parent_path()),
          filename_(output_filename.stem().string()),
          extension_(output_filename.extension().string()),
          ascii_(ascii) {
        if (extension_ != ".vtu" && extension_ != ".pvtu")
            throw std::invalid_argument("DataWriter_t: the output extension must be .vtu or .pvtu, got '" +
                                        extension_ + "'");
        series_filename_ = output_filename;
        series_filename_ += ".series";
        series_element_filename_ = directory_ / filename_;
        series_element_filename_ += "_element" + extension_ + ".series";
        if (!directory_.empty()) std::filesystem::create_directories(directory_);
        create_time_series_file();
    }

    // DataWriter_t::write_data (DataWriter_t.cpp:36-133).
    void write_data(std::size_t n_interpolation_points, std::size_t n_elements, double time,
                    const std::vector<double> &x, const std::vector<double> &y, const std::vector<double> &p,
                    const std::vector<double> &u, const std::vector<double> &v) const {
        const std::string name = filename_ + time_string(time) + extension_;
        std::vector<detail::Array> point_data{detail::make_array("Pressure", p, 1, ascii_),
                                              detail::make_array("Velocity", detail::interleave(u, v), 2, ascii_)};
        write_grid(name, points(x, y), connectivity(n_interpolation_points, n_elements), point_data, {});
        series_.emplace_back(name, time);
        write_series(series_filename_, series_);
    }

    // DataWriter_t::write_complete_data (DataWriter_t.cpp:135-503).
    void write_complete_data(std::size_t n_interpolation_points, std::size_t n_elements, double time, int rank,
                             const std::vector<double> &x, const std::vector<double> &y,
                             const std::vector<double> &p, const std::vector<double> &u,
                             const std::vector<double> &v, const std::vector<int> &N,
                             const std::vector<double> &dp_dt, const std::vector<double> &du_dt,
                             const std::vector<double> &dv_dt, const std::vector<double> &p_error,
                             const std::vector<double> &u_error, const std::vector<double> &v_error,
                             const std::vector<double> &p_sigma, const std::vector<double> &u_sigma,
                             const std::vector<double> &v_sigma, const std::vector<int> &refine,
                             const std::vector<int> &coarsen, const std::vector<int> &split_level,
                             const std::vector<double> &p_analytical_error,
                             const std::vector<double> &u_analytical_error,
                             const std::vector<double> &v_analytical_error, const std::vector<int> &status,
                             const std::vector<int> &rotation) const {
        using detail::interleave;
        using detail::make_array;
        const std::size_t n = n_interpolation_points, K = n_elements;
        const std::string ts = time_string(time);
        const std::string name = filename_ + ts + extension_;
        const std::string name_element = filename_ + "_element" + ts + extension_;
        std::vector<detail::Array> point_data{
            make_array("Pressure", p, 1, ascii_),
            make_array("Velocity", interleave(u, v), 2, ascii_),
            make_array("PressureDerivative", dp_dt, 1, ascii_),
            make_array("VelocityDerivative", interleave(du_dt, dv_dt), 2, ascii_),
            make_array("PressureAnalyticalError", p_analytical_error, 1, ascii_),
            make_array("VelocityAnalyticalError", interleave(u_analytical_error, v_analytical_error), 2, ascii_)};
        std::vector<uint64_t> index(K);
        for (std::size_t e = 0; e < K; ++e) index[e] = e;
        auto i32 = [](const std::vector<int> &a) { return std::vector<int32_t>(a.begin(), a.end()); };
        std::vector<detail::Array> cell_data{
            make_array("N", i32(N), 1, ascii_),
            make_array("Index", index, 1, ascii_),
            make_array("Status", i32(status), 1, ascii_),
            make_array("Rotation", i32(rotation), 1, ascii_),
            make_array("Rank", std::vector<int32_t>(K, rank), 1, ascii_),
            make_array("PressureError", p_error, 1, ascii_),
            make_array("VelocityError", interleave(u_error, v_error), 2, ascii_),
            make_array("PressureSigma", p_sigma, 1, ascii_),
            make_array("VelocitySigma", interleave(u_sigma, v_sigma), 2, ascii_),
            make_array("Refine", i32(refine), 1, ascii_),
            make_array("Coarsen", i32(coarsen), 1, ascii_),
            make_array("SplitLevel", i32(split_level), 1, ascii_)};
        write_grid(name, points(x, y), connectivity(n, K), point_data, {});
        // element grid: output points (0, 0), (n-1, 0), (n-1, n-1), (0, n-1) (DataWriter_t.cpp:175-184)
        std::vector<double> corners(12 * K, 0.0);
        const std::size_t sel[4] = {0, (n - 1) * n, (n - 1) * n + n - 1, n - 1};
        for (std::size_t e = 0; e < K; ++e)
            for (int c = 0; c < 4; ++c) {
                corners[3 * (4 * e + c)] = x[e * n * n + sel[c]];
                corners[3 * (4 * e + c) + 1] = y[e * n * n + sel[c]];
            }
        std::vector<int64_t> cells(4 * K);
        for (std::size_t i = 0; i < 4 * K; ++i) cells[i] = (int64_t)i;
        write_grid(name_element, corners, cells, {}, cell_data);
        series_.emplace_back(name, time);
        series_element_.emplace_back(name_element, time);
        write_series(series_filename_, series_);
        write_series(series_element_filename_, series_element_);
    }

    // "_t<time>s": std::fixed, 9 decimals, '.' replaced by '_' (DataWriter_t.cpp:113-116).
    static std::string time_string(double time) {
        std::ostringstream ss;
        ss << "_t" << std::setprecision(9) << std::fixed << time << "s";
        std::string s = ss.str();
        std::replace(s.begin(), s.end(), '.', '_');
        return s;
    }

    std::filesystem::path directory_;
    std::string filename_;
    std::string extension_;
    std::filesystem::path series_filename_;
    std::filesystem::path series_element_filename_;

  private:
    using Series = std::vector<std::pair<std::string, double>>;

    void create_time_series_file() const {
        series_.clear();
        series_element_.clear();
        write_series(series_filename_, series_);
        write_series(series_element_filename_, series_element_);
    }

    static void write_series(const std::filesystem::path &path, const Series &files) {
        std::ofstream o(path);
        o << "{\n  \"file-series-version\": \"1.0\",\n  \"files\": ";
        if (files.empty()) {
            o << "[]\n}\n";
        } else {
            o << "[\n";
            for (std::size_t i = 0; i < files.size(); ++i)
                o << "    {\n      \"name\": \"" << files[i].first << "\",\n      \"time\": "
                  << detail::number(files[i].second, true) << "\n    }" << (i + 1 < files.size() ? ",\n" : "\n");
            o << "  ]\n}\n";
        }
        if (!o) throw std::runtime_error("DataWriter_t: could not write " + path.string());
    }

    static std::vector<double> points(const std::vector<double> &x, const std::vector<double> &y) {
        std::vector<double> pts(3 * x.size(), 0.0);
        for (std::size_t i = 0; i < x.size(); ++i) {
            pts[3 * i] = x[i];
            pts[3 * i + 1] = y[i];
        }
        return pts;
    }

    static std::vector<int64_t> connectivity(std::size_t n, std::size_t K) {
        std::vector<int64_t> c;
        c.reserve(4 * K * (n - 1) * (n - 1));
        for (std::size_t e = 0; e < K; ++e) {
            const int64_t o = (int64_t)(e * n * n), m = (int64_t)n;
            for (int64_t i = 0; i + 1 < m; ++i)
                for (int64_t j = 0; j + 1 < m; ++j) {
                    c.push_back(o + (i + 1) * m + j);
                    c.push_back(o + (i + 1) * m + j + 1);
                    c.push_back(o + i * m + j + 1);
                    c.push_back(o + i * m + j);
                }
        }
        return c;
    }

    void write_vtu(const std::filesystem::path &path, const std::vector<double> &pts, const std::vector<int64_t> &cells,
                   const std::vector<detail::Array> &point_data, const std::vector<detail::Array> &cell_data) const {
        const std::size_t n_cells = cells.size() / 4;
        std::vector<int64_t> offsets(n_cells);
        for (std::size_t c = 0; c < n_cells; ++c) offsets[c] = 4 * (int64_t)(c + 1);
        std::ofstream o(path);
        o << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"UnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
          << "  <UnstructuredGrid>\n"
          << "    <Piece NumberOfPoints=\"" << pts.size() / 3 << "\" NumberOfCells=\"" << n_cells << "\">\n"
          << "      <PointData>\n";
        for (const auto &a : point_data) detail::data_array(o, a, ascii_);
        o << "      </PointData>\n      <CellData>\n";
        for (const auto &a : cell_data) detail::data_array(o, a, ascii_);
        o << "      </CellData>\n      <Points>\n";
        detail::data_array(o, detail::make_array("Points", pts, 3, ascii_), ascii_);
        o << "      </Points>\n      <Cells>\n";
        detail::data_array(o, detail::make_array("connectivity", cells, 1, ascii_), ascii_);
        detail::data_array(o, detail::make_array("offsets", offsets, 1, ascii_), ascii_);
        detail::data_array(o, detail::make_array("types", std::vector<uint8_t>(n_cells, 9), 1, ascii_), ascii_);
        o << "      </Cells>\n    </Piece>\n  </UnstructuredGrid>\n</VTKFile>\n";
        if (!o) throw std::runtime_error("DataWriter_t: could not write " + path.string());
    }

    void write_grid(const std::string &name, const std::vector<double> &pts, const std::vector<int64_t> &cells,
                    const std::vector<detail::Array> &point_data, const std::vector<detail::Array> &cell_data) const {
        if (extension_ != ".pvtu") {
            write_vtu(directory_ / name, pts, cells, point_data, cell_data);
            return;
        }
        const std::string piece = std::filesystem::path(name).stem().string() + "_0.vtu";
        write_vtu(directory_ / piece, pts, cells, point_data, cell_data);
        std::ofstream o(directory_ / name);
        o << "<?xml version=\"1.0\"?>\n"
          << "<VTKFile type=\"PUnstructuredGrid\" version=\"1.0\" byte_order=\"LittleEndian\" header_type=\"UInt64\">\n"
          << "  <PUnstructuredGrid GhostLevel=\"0\">\n    <PPointData>\n";
        for (const auto &a : point_data) o << "      <PDataArray " << detail::attrs(a) << "/>\n";
        o << "    </PPointData>\n    <PCellData>\n";
        for (const auto &a : cell_data) o << "      <PDataArray " << detail::attrs(a) << "/>\n";
        o << "    </PCellData>\n    <PPoints>\n"
          << "      <PDataArray type=\"Float64\" Name=\"Points\" NumberOfComponents=\"3\"/>\n"
          << "    </PPoints>\n    <Piece Source=\"" << piece << "\"/>\n  </PUnstructuredGrid>\n</VTKFile>\n";
        if (!o) throw std::runtime_error("DataWriter_t: could not write " + (directory_ / name).string());
    }

    bool ascii_;
    mutable Series series_, series_element_;
};

}  // namespace Helpers
}  // namespace SEM

#endif
