// test_image_io.cpp -- include/ndl/image_io.h on the host ("label: Pass/Fail"
// lines, exit status = number of failures).  Built and run by
// tests/test_image_io.py with a scratch directory as argv[1]; that test then
// compares the files written here with the reference's own NRRD writer and
// reads them with the reference's loader.
//
// Every type's pattern value(i) (i = dense index, dimension 0 fastest) is
// also defined in tests/test_image_io.py.
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include <ndl/image.h>
#include <ndl/image_io.h>

using namespace ndl;

static int failures = 0;
static void report(const std::string& label, bool ok)
{
    std::printf("%s: %s\n", label.c_str(), ok ? "Pass" : "Fail");
    std::fflush(stdout);
    if (!ok) failures++;
}

template <class T> T pattern(long i);
template <> float pattern<float>(long i) { return 0.25f * (float)i - 3.0f; }
template <> unsigned char pattern<unsigned char>(long i) { return (unsigned char)((i * 7) % 256); }
template <> short pattern<short>(long i) { return (short)(i * 37 - 300); }
template <> unsigned short pattern<unsigned short>(long i) { return (unsigned short)((i * 1000) % 65536); }
template <> int pattern<int>(long i) { return (int)(i * 100003 - 7); }
template <> unsigned int pattern<unsigned int>(long i) { return (unsigned int)((unsigned long long)i * 123456789ull); }
template <> double pattern<double>(long i) { return (double)i * 1e-3 + 1.0 / 3.0; }

template <class T, int DIM> OwnedImage<T, DIM> make(std::array<int, DIM> e)
{
    OwnedImage<T, DIM> img(e);
    long i = 0;
    for (auto it = img.begin(); it != img.end(); ++it) *it = pattern<T>(i++);
    return img;
}

template <class T, int DIM> void round_trip(const std::string& dir, const std::string& tag, std::array<int, DIM> e)
{
    OwnedImage<T, DIM> img = make<T, DIM>(e);
    const std::string path = dir + "/ours_" + tag + ".nrrd";
    image_io::save(img, path);
    std::array<int, DIM> got{};
    std::vector<T> v = image_io::load_nrrd<T, DIM>(got, path);
    bool ok = got == e && v.size() == img.size();
    for (std::size_t i = 0; ok && i < v.size(); i++) ok = v[i] == pattern<T>((long)i);
    report("NRRD " + tag + " round trip (same type)", ok);
    std::vector<double> d = image_io::load_nrrd<double, DIM>(got, path);
    ok = d.size() == img.size();
    for (std::size_t i = 0; ok && i < d.size(); i++) ok = d[i] == (double)pattern<T>((long)i);
    report("NRRD " + tag + " loads converted to double", ok);
}

static void write_file(const std::string& path, const std::string& header, const std::vector<unsigned char>& data)
{
    std::ofstream f(path, std::ios::binary);
    f.write(header.data(), (std::streamsize)header.size());
    f.write((const char*)data.data(), (std::streamsize)data.size());
}
template <class T> static void put(std::vector<unsigned char>& b, T v, bool big)
{
    unsigned char c[sizeof(T)];
    std::memcpy(c, &v, sizeof(T));
    if (big) std::reverse(c, c + sizeof(T));
    b.insert(b.end(), c, c + sizeof(T));
}
template <class F> static bool throws(F&& f)
{
    try {
        f();
    } catch (const std::runtime_error&) {
        return true;
    }
    return false;
}

int main(int argc, char** argv)
{
    if (argc < 2) {
        std::printf("usage: test_image_io <scratch dir>\n");
        return 1;
    }
    const std::string dir = argv[1];

    round_trip<float, 3>(dir, "f32", {5, 4, 3});
    round_trip<unsigned char, 2>(dir, "u8", {7, 6});
    round_trip<short, 3>(dir, "i16", {9, 2, 2});
    round_trip<unsigned short, 2>(dir, "u16", {3, 3});
    round_trip<int, 2>(dir, "i32", {4, 5});
    round_trip<unsigned int, 1>(dir, "u32", {6});
    round_trip<double, 4>(dir, "f64", {2, 3, 4, 2});

    // a strided view is written densely in its own index order
    {
        OwnedImage<float, 3> img = make<float, 3>({5, 4, 3});
        auto mirrored = img.view({}, {}, {1, 1, -1}); // z reversed
        image_io::save(mirrored, dir + "/ours_f32_mirror.nrrd");
        std::array<int, 3> e{};
        std::vector<float> v = image_io::load_nrrd<float, 3>(e, dir + "/ours_f32_mirror.nrrd");
        bool ok = e == img.extent();
        for (int z = 0; ok && z < 3; z++)
            for (int y = 0; y < 4; y++)
                for (int x = 0; x < 5; x++) ok = ok && v[(z * 4 + y) * 5 + x] == img(x, y, 2 - z);
        report("NRRD save of a mirrored view writes it densely", ok);
        // no extension: ".nrrd" is appended (imageIO.h:260-265)
        image_io::save(mirrored, dir + "/ours_noext");
        std::ifstream f(dir + "/ours_noext.nrrd", std::ios::binary);
        report("NRRD save without extension appends .nrrd", (bool)f);
        report("save with another extension throws",
               throws([&] { image_io::save(mirrored, dir + "/x.png"); }));
    }

    // raw files
    {
        OwnedImage<float, 3> img = make<float, 3>({5, 4, 3});
        image_io::save_raw(img, dir + "/ours.raw");
        std::vector<float> v = image_io::load_raw<float, 3>(dir + "/ours.raw", {5, 4, 3}, 0);
        bool ok = v.size() == 60;
        for (int i = 0; ok && i < 60; i++) ok = v[i] == pattern<float>(i);
        report("raw round trip", ok);
        std::vector<float> s = image_io::load_raw<float, 3>(dir + "/ours.raw", {5, 4, 3}, 8);
        ok = s.size() == 60;
        for (int i = 0; ok && i < 58; i++) ok = s[i] == pattern<float>(i + 2);
        ok = ok && s[58] == 0.f && s[59] == 0.f;
        report("raw load at a byte offset; a short file leaves zeros (imageIO.h:146-157)", ok);
        std::vector<float> m = image_io::load_raw<float, 3>(dir + "/missing.raw", {2, 2, 1}, 0);
        report("raw load of a missing file gives zeros", m.size() == 4 && m[0] == 0.f && m[3] == 0.f);
    }

    // header variants the reference's loader accepts
    {
        std::vector<unsigned char> b;
        for (int i = 0; i < 6; i++) put<float>(b, 0.5f * (float)i + 1.0f, true);
        write_file(dir + "/v_big.nrrd",
                   "NRRD0004\ntype: float\ndimension: 2\nsizes: 3 2\nencoding: raw\nendian: big\n\n", b);
        std::array<int, 2> e{};
        std::vector<float> v = image_io::load_nrrd<float, 2>(e, dir + "/v_big.nrrd");
        bool ok = e == std::array<int, 2>{3, 2};
        for (int i = 0; ok && i < 6; i++) ok = v[i] == 0.5f * (float)i + 1.0f;
        report("NRRD big-endian data is byte-swapped", ok);
    }
    {
        write_file(dir + "/v_comments.nrrd",
                   "NRRD0005\n# a comment line\ntype: uint8\ndimension: 1\nsizes: 4\nencoding: raw\nendian: little\n"
                   "spacing: 1.5\nmykey:=some value\n\n",
                   {10, 20, 30, 40});
        std::array<int, 1> e{};
        std::vector<int> v = image_io::load_nrrd<int, 1>(e, dir + "/v_comments.nrrd");
        report("NRRD comments, key:=value pairs and type alias uint8",
               e[0] == 4 && v == std::vector<int>{10, 20, 30, 40});
    }
    {
        std::vector<unsigned char> b = {1, 2, 3, 4, 5, 6};
        for (int i = 0; i < 4; i++) put<short>(b, (short)(-3 * i), false);
        write_file(dir + "/v_skip.nrrd",
                   "NRRD0004\ntype: int16\ndimension: 2\nsizes: 2 2\nencoding: raw\nendian: little\nbyte skip: 6\n\n", b);
        std::array<int, 2> e{};
        std::vector<double> v = image_io::load_nrrd<double, 2>(e, dir + "/v_skip.nrrd");
        report("NRRD byte skip", v == std::vector<double>{0, -3, -6, -9});
    }
    {
        std::vector<unsigned char> b;
        for (int i = 0; i < 3; i++) put<unsigned short>(b, (unsigned short)(1000 + i), false);
        write_file(dir + "/v_data.raw", "", b);
        write_file(dir + "/v_datafile.nrrd",
                   "NRRD0004\ntype: unsigned short\ndimension: 1\nsizes: 3\nencoding: raw\nendian: little\ndatafile: " +
                       dir + "/v_data.raw\n\n",
                   {});
        std::array<int, 1> e{};
        std::vector<unsigned short> v = image_io::load_nrrd<unsigned short, 1>(e, dir + "/v_datafile.nrrd");
        report("NRRD external data file", v == std::vector<unsigned short>{1000, 1001, 1002});
    }

    // errors (the reference prints them and leaves the data undefined)
    write_file(dir + "/v_gzip.nrrd", "NRRD0004\ntype: float\ndimension: 1\nsizes: 2\nencoding: gzip\n\n", {0, 0});
    write_file(dir + "/v_lineskip.nrrd",
               "NRRD0004\ntype: uint8\ndimension: 1\nsizes: 2\nencoding: raw\nline skip: 1\n\n", {1, 2});
    write_file(dir + "/v_short.nrrd", "NRRD0004\ntype: float\ndimension: 1\nsizes: 4\nencoding: raw\n\n", {0, 0, 0});
    write_file(dir + "/v_type.nrrd", "NRRD0004\ntype: block\ndimension: 1\nsizes: 4\nencoding: raw\n\n", {0, 0, 0, 0});
    write_file(dir + "/v_magic.nrrd", "P5\n4 1\n255\n", {1, 2, 3, 4});
    std::array<int, 1> e1{};
    std::array<int, 2> e2{};
    report("NRRD gzip encoding throws", throws([&] { image_io::load_nrrd<float, 1>(e1, dir + "/v_gzip.nrrd"); }));
    report("NRRD line skip throws", throws([&] { image_io::load_nrrd<float, 1>(e1, dir + "/v_lineskip.nrrd"); }));
    report("NRRD short data throws", throws([&] { image_io::load_nrrd<float, 1>(e1, dir + "/v_short.nrrd"); }));
    report("NRRD unknown type throws", throws([&] { image_io::load_nrrd<float, 1>(e1, dir + "/v_type.nrrd"); }));
    report("NRRD bad magic throws", throws([&] { image_io::load_nrrd<float, 1>(e1, dir + "/v_magic.nrrd"); }));
    report("NRRD missing file throws", throws([&] { image_io::load_nrrd<float, 1>(e1, dir + "/nothing.nrrd"); }));
    report("NRRD dimension mismatch throws",
           throws([&] { image_io::load_nrrd<float, 2>(e2, dir + "/ours_f32.nrrd"); }));
    return failures;
}
