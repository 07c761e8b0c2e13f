// image_io.h -- headerless raw and NRRD files for the images of the
// convolution path, on the host and straight into device images.
//
// Host functions keep the names and argument meaning of the reference's
// image_io namespace (imageIO.h:146-170 load_raw / save_raw, :254-277 save
// for .nrrd or no extension, :383-398 load_nrrd) and its NRRD format
// (imageIO/NRRD/nrrd.h:114-164 save, :166-231 header parsing and type names,
// :233-366 load): files written here are byte-identical to the reference's,
// and every file the reference writes or reads loads here with the same
// values.  The device functions (no counterpart in the reference) read the
// file into pinned host memory (ndl_host_alloc) and upload it with one
// asynchronous copy into a DeviceImage, the entry point for 3-D/4-D data
// that the convolution then keeps in HBM.
//
// Errors: the reference reports NRRD failures on stderr and leaves the data
// undefined; here they throw std::runtime_error (bad magic or header line,
// unsupported type / encoding / "line skip", short data, missing file).
// load_raw keeps the reference's behaviour: a short or missing file leaves the
// remaining elements zero.
#pragma once
#include <algorithm>
#include <array>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <functional>
#include <iomanip>
#include <map>
#include <numeric>
#include <sstream>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "image.h"

namespace ndl {
namespace image_io {
namespace detail {

// The "type" field the reference writes: its C type name (nrrd.h:123-127
// writes typeName<T>(), with char -> "int8" and unsigned int -> "uint32").
template <class T> inline const char* nrrd_type_name()
{
    if constexpr (std::is_same_v<T, char>) return "int8";
    else if constexpr (std::is_same_v<T, signed char>) return "int8";
    else if constexpr (std::is_same_v<T, unsigned char>) return "unsigned char";
    else if constexpr (std::is_same_v<T, short>) return "short";
    else if constexpr (std::is_same_v<T, unsigned short>) return "unsigned short";
    else if constexpr (std::is_same_v<T, int>) return "int";
    else if constexpr (std::is_same_v<T, unsigned int>) return "uint32";
    else if constexpr (std::is_same_v<T, float>) return "float";
    else if constexpr (std::is_same_v<T, double>) return "double";
    else static_assert(sizeof(T) == 0, "image_io: no NRRD type for this element type");
}

// Header type strings the reference accepts, mapped to one C type each
// (nrrd.h:209-231).
enum class FileType { I8, U8, I16, U16, I32, U32, F32, F64 };
inline FileType nrrd_file_type(const std::string& t)
{
    static const std::map<std::string, FileType> m = {
        {"int8_t", FileType::I8}, {"int8", FileType::I8}, {"signed char", FileType::I8}, {"char", FileType::I8},
        {"uchar", FileType::U8}, {"uint8_t", FileType::U8}, {"uint8", FileType::U8},
        {"unsigned char", FileType::U8},
        {"int16_t", FileType::I16}, {"int16", FileType::I16}, {"signed short int", FileType::I16},
        {"signed signed", FileType::I16}, {"short int", FileType::I16}, {"short", FileType::I16},
        {"ushort", FileType::U16}, {"uint16_t", FileType::U16}, {"uint16", FileType::U16},
        {"unsigned short int", FileType::U16}, {"unsigned short", FileType::U16},
        {"int32_t", FileType::I32}, {"int32", FileType::I32}, {"signed int", FileType::I32}, {"int", FileType::I32},
        {"uint32_t", FileType::U32}, {"uint32", FileType::U32}, {"uint", FileType::U32},
        {"unsigned int", FileType::U32},
        {"float", FileType::F32}, {"double", FileType::F64},
    };
    auto it = m.find(t);
    if (it == m.end()) throw std::runtime_error("NRRD: unsupported type '" + t + "'");
    return it->second;
}
inline std::size_t file_type_size(FileType t)
{
    switch (t) {
    case FileType::I8: case FileType::U8: return 1;
    case FileType::I16: case FileType::U16: return 2;
    case FileType::I32: case FileType::U32: case FileType::F32: return 4;
    case FileType::F64: return 8;
    }
    return 0;
}
template <class T> constexpr bool same_file_type(FileType t)
{
    switch (t) {
    case FileType::I8: return std::is_same_v<T, signed char> || std::is_same_v<T, char>;
    case FileType::U8: return std::is_same_v<T, unsigned char>;
    case FileType::I16: return std::is_same_v<T, short>;
    case FileType::U16: return std::is_same_v<T, unsigned short>;
    case FileType::I32: return std::is_same_v<T, int>;
    case FileType::U32: return std::is_same_v<T, unsigned int>;
    case FileType::F32: return std::is_same_v<T, float>;
    case FileType::F64: return std::is_same_v<T, double>;
    }
    return false;
}

inline bool host_big_endian()
{
    const uint16_t one = 1;
    unsigned char b;
    std::memcpy(&b, &one, 1);
    return b == 0;
}

// Raw element bytes of type `ft` (byte-swapped when `swap`) converted to T
// with a plain cast per element, as the reference's convert<X,T> does.
template <class T> void convert_elements(const unsigned char* raw, FileType ft, bool swap, std::size_t n, T* out)
{
    auto run = [&](auto tag) {
        using S = decltype(tag);
        for (std::size_t i = 0; i < n; i++) {
            unsigned char b[sizeof(S)];
            std::memcpy(b, raw + i * sizeof(S), sizeof(S));
            if (swap) std::reverse(b, b + sizeof(S));
            S v;
            std::memcpy(&v, b, sizeof(S));
            out[i] = (T)v;
        }
    };
    switch (ft) {
    case FileType::I8: run((signed char)0); break;
    case FileType::U8: run((unsigned char)0); break;
    case FileType::I16: run((short)0); break;
    case FileType::U16: run((unsigned short)0); break;
    case FileType::I32: run((int)0); break;
    case FileType::U32: run((unsigned int)0); break;
    case FileType::F32: run(0.0f); break;
    case FileType::F64: run(0.0); break;
    }
}

// A parsed header: where the data is and how it is laid out.
struct NrrdInfo {
    std::map<std::string, std::string> fields; // "key: value" lines
    std::map<std::string, std::string> keys;   // "key:=value" lines
    std::vector<int> sizes;
    FileType type = FileType::U8;
    bool swap = false;
    std::string data_file;
    std::size_t offset = 0; // byte offset of the data in data_file
    std::size_t count() const
    {
        std::size_t n = 1;
        for (int s : sizes) n *= (std::size_t)s;
        return n;
    }
};

// Header grammar of nrrd.h:166-207: magic line, then until the first empty
// line, '#' comments, "key: value" fields and "key:=value" pairs.
inline NrrdInfo read_nrrd_header(const std::string& path)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("NRRD: cannot open " + path);
    std::string line;
    if (!std::getline(f, line) || line.compare(0, 4, "NRRD") != 0)
        throw std::runtime_error("NRRD: " + path + " does not start with a NRRD magic line");
    NrrdInfo info;
    bool ended = false;
    while (std::getline(f, line)) {
        if (line.empty()) {
            ended = true;
            break;
        }
        if (line[0] == '#') continue;
        const std::size_t c = line.find(':');
        if (c == std::string::npos || c + 1 >= line.size())
            throw std::runtime_error("NRRD: cannot parse header line '" + line + "' in " + path);
        if (line[c + 1] == ' ') info.fields[line.substr(0, c)] = line.substr(c + 2);
        else if (line[c + 1] == '=') info.keys[line.substr(0, c)] = line.substr(c + 2);
        else throw std::runtime_error("NRRD: cannot parse header line '" + line + "' in " + path);
    }
    if (!ended) throw std::runtime_error("NRRD: header of " + path + " has no terminating empty line");
    info.offset = (std::size_t)f.tellg();
    auto field = [&](const char* a, const char* b = nullptr) -> const std::string* {
        auto it = info.fields.find(a);
        if (it == info.fields.end() && b) it = info.fields.find(b);
        return it == info.fields.end() ? nullptr : &it->second;
    };
    const std::string* t = field("type");
    if (!t) throw std::runtime_error("NRRD: no type field in " + path);
    info.type = nrrd_file_type(*t);
    const std::string* enc = field("encoding");
    if (!enc || *enc != "raw") throw std::runtime_error("NRRD: only raw encoding is supported (" + path + ")");
    // the reference always writes an endian field and reads a missing one as little
    const std::string* en = field("endian");
    const bool file_big = en && *en != "little";
    info.swap = file_big != host_big_endian();
    const std::string* dim = field("dimension");
    const std::string* sz = field("sizes");
    int n = dim ? std::atoi(dim->c_str()) : 0;
    if (sz) {
        std::istringstream ss(*sz);
        int v;
        while (ss >> v) info.sizes.push_back(v);
    }
    if (n <= 0 || n > 10 || (int)info.sizes.size() != n)
        throw std::runtime_error("NRRD: bad dimension / sizes in " + path);
    for (int s : info.sizes)
        if (s <= 0) throw std::runtime_error("NRRD: bad sizes in " + path);
    info.data_file = path;
    if (const std::string* df = field("datafile", "data file")) {
        info.data_file = *df;
        info.offset = 0;
    }
    if (const std::string* bs = field("byteskip", "byte skip")) {
        const long skip = std::atol(bs->c_str());
        if (skip < 0) throw std::runtime_error("NRRD: negative byte skip is not supported (" + path + ")");
        info.offset += (std::size_t)skip;
    }
    if (field("lineskip", "line skip")) throw std::runtime_error("NRRD: line skip is not supported (" + path + ")");
    return info;
}

inline void read_bytes(const std::string& path, std::size_t offset, void* dst, std::size_t bytes)
{
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("NRRD: cannot open data file " + path);
    f.seekg((std::streamoff)offset, std::ios::beg);
    f.read((char*)dst, (std::streamsize)bytes);
    if (!f || (std::size_t)f.gcount() != bytes) throw std::runtime_error("NRRD: failed to read the complete data of " + path);
}

// Header of a file the reference would write (nrrd.h:114-164): magic,
// fields in key order, then a comment giving the data offset (8 characters,
// right aligned), an empty line, and the dense data.
template <class T> std::string nrrd_header(const std::vector<int>& sizes)
{
    std::map<std::string, std::string> f;
    f["type"] = nrrd_type_name<T>();
    f["dimension"] = std::to_string(sizes.size());
    std::string s;
    for (std::size_t i = 0; i < sizes.size(); i++) s += (i ? " " : "") + std::to_string(sizes[i]);
    f["sizes"] = s;
    f["encoding"] = "raw";
    f["endian"] = host_big_endian() ? "big" : "little";
    std::string h = "NRRD0004\n";
    for (const auto& kv : f) h += kv.first + ": " + kv.second + "\n";
    h += "# Offset to raw data: ";
    const std::string tail = " bytes.\n\n";
    std::ostringstream off;
    off << std::setw(8) << std::setfill(' ') << (h.size() + tail.size() + 8);
    return h + off.str() + tail;
}

// Pinned host memory from the library (page-locked: one DMA copy to HBM).
struct PinnedBuffer {
    void* p = nullptr;
    explicit PinnedBuffer(std::size_t bytes) { ndl::detail::check(ndl_host_alloc(&p, bytes ? bytes : 1)); }
    ~PinnedBuffer()
    {
        if (p) (void)ndl_host_free(p);
    }
    PinnedBuffer(const PinnedBuffer&) = delete;
    PinnedBuffer& operator=(const PinnedBuffer&) = delete;
};

template <class T, int DIM> std::array<int, DIM> extent_of(const NrrdInfo& info, const std::string& path)
{
    if ((int)info.sizes.size() != DIM)
        throw std::runtime_error("NRRD: " + path + " has " + std::to_string(info.sizes.size()) +
                                 " dimensions, the image type " + std::to_string(DIM));
    std::array<int, DIM> e{};
    std::copy(info.sizes.begin(), info.sizes.end(), e.begin());
    return e;
}

} // namespace detail

/// Loads a headerless raw file of `extent` elements of T starting at
/// `offsetBytes`, in the file's byte order (imageIO.h:146-157).
template <class T, int DIM>
std::vector<T> load_raw(std::string rawFileName, std::array<int, DIM> extent, int offsetBytes)
{
    const std::size_t n = Image<T, DIM>::size(extent);
    std::vector<T> result(n);
    std::ifstream is(rawFileName, std::ios::binary);
    is.seekg(offsetBytes, std::ios::beg);
    is.read((char*)result.data(), (std::streamsize)(n * sizeof(T)));
    return result;
}

/// Writes the image's elements, dimension 0 fastest, with no header
/// (imageIO.h:159-170).
template <class T, int DIM> void save_raw(const Image<T, DIM>& source, std::string rawFileName)
{
    OwnedImage<T, DIM> dense(source.extent());
    dense.assign(source);
    std::ofstream f(rawFileName, std::ios::binary);
    if (!f) throw std::runtime_error("save_raw: cannot write " + rawFileName);
    f.write((const char*)dense.data(), (std::streamsize)(dense.size() * sizeof(T)));
}

/// Saves as NRRD (".nrrd", or no extension: ".nrrd" is appended), the
/// reference's format byte for byte (imageIO.h:254-277).  Other extensions
/// (bitmap / PNG codecs) are not part of this library.
template <class T, int DIM> void save(const Image<T, DIM>& image, std::string fileName)
{
    const std::size_t dot = fileName.find_last_of('.');
    std::string ext = dot == std::string::npos ? "" : fileName.substr(dot + 1);
    if (ext.empty()) fileName += ".nrrd";
    else if (ext != "nrrd") throw std::runtime_error("image_io::save: only .nrrd is supported (" + fileName + ")");
    OwnedImage<T, DIM> dense(image.extent());
    dense.assign(image);
    const std::string h = detail::nrrd_header<T>(std::vector<int>(image.extent().begin(), image.extent().end()));
    std::ofstream f(fileName, std::ios::binary);
    if (!f) throw std::runtime_error("image_io::save: cannot write " + fileName);
    f.write(h.data(), (std::streamsize)h.size());
    f.write((const char*)dense.data(), (std::streamsize)(dense.size() * sizeof(T)));
}

/// Loads a NRRD file converted to T; `extent` receives its sizes
/// (imageIO.h:383-398, nrrd.h:233-366).
template <class T, int DIM> std::vector<T> load_nrrd(std::array<int, DIM>& extent, std::string fileName)
{
    const detail::NrrdInfo info = detail::read_nrrd_header(fileName);
    extent = detail::extent_of<T, DIM>(info, fileName);
    const std::size_t n = info.count(), fb = detail::file_type_size(info.type);
    std::vector<T> out(n);
    if (detail::same_file_type<T>(info.type) && !info.swap) {
        detail::read_bytes(info.data_file, info.offset, out.data(), n * sizeof(T));
    } else {
        std::vector<unsigned char> raw(n * fb);
        detail::read_bytes(info.data_file, info.offset, raw.data(), raw.size());
        detail::convert_elements(raw.data(), info.type, info.swap, n, out.data());
    }
    return out;
}

/// Reads a raw file into pinned host memory and uploads it into a new
/// device image (no counterpart in the reference).  Unlike load_raw, a short
/// file is an error.
template <class T, int DIM>
DeviceImage<T, DIM> load_raw_device(const std::string& rawFileName, std::array<int, DIM> extent, int offsetBytes,
                                    void* stream = nullptr)
{
    DeviceImage<T, DIM> img(extent);
    const std::size_t bytes = img.size() * sizeof(T);
    detail::PinnedBuffer pin(bytes);
    detail::read_bytes(rawFileName, (std::size_t)offsetBytes, pin.p, bytes);
    ndl::detail::check(ndl_memcpy_htod(img.data(), pin.p, bytes, stream));
    ndl::detail::check(ndl_stream_synchronize(stream));
    return img;
}

/// Loads a NRRD file into a new device image of T through pinned host
/// memory: the data is read straight into the pinned buffer (converted there
/// when the file type or byte order differs) and uploaded with one copy.
template <class T, int DIM> DeviceImage<T, DIM> load_nrrd_device(const std::string& fileName, void* stream = nullptr)
{
    const detail::NrrdInfo info = detail::read_nrrd_header(fileName);
    DeviceImage<T, DIM> img(detail::extent_of<T, DIM>(info, fileName));
    const std::size_t n = info.count(), bytes = n * sizeof(T);
    detail::PinnedBuffer pin(bytes);
    if (detail::same_file_type<T>(info.type) && !info.swap) {
        detail::read_bytes(info.data_file, info.offset, pin.p, bytes);
    } else {
        std::vector<unsigned char> raw(n * detail::file_type_size(info.type));
        detail::read_bytes(info.data_file, info.offset, raw.data(), raw.size());
        detail::convert_elements(raw.data(), info.type, info.swap, n, (T*)pin.p);
    }
    ndl::detail::check(ndl_memcpy_htod(img.data(), pin.p, bytes, stream));
    ndl::detail::check(ndl_stream_synchronize(stream));
    return img;
}

/// Saves a device image as NRRD (the format of save()), downloading through
/// pinned host memory.
template <class T, int DIM> void save(const DeviceImage<T, DIM>& image, std::string fileName, void* stream = nullptr)
{
    const std::size_t bytes = image.size() * sizeof(T);
    detail::PinnedBuffer pin(bytes);
    ndl::detail::check(ndl_stream_synchronize(stream));
    ndl::detail::check(ndl_memcpy_dtoh(pin.p, image.data(), bytes, stream));
    ndl::detail::check(ndl_stream_synchronize(stream));
    save(Image<T, DIM>((T*)pin.p, image.extent()), fileName);
}

} // namespace image_io
} // namespace ndl
