// test_dropin.cpp -- the drop-in C++ API (include/ndl) on the GPU, written the
// way the reference's own unit tests are ("label: Pass/Fail" lines,
// unitTests/testHelpers.h:112-120).  Built by __graft_entry__.build(); run by
// tests/test_gpu_dropin.py on the GPU.  Exit status = number of failures.
#include <cmath>
#include <cstdio>
#include <random>
#include <vector>

#include <ndl/convolution.h>
#include <ndl/image.h>
#include <ndl/image_io.h>
#include <ndl/morphology.h>

using namespace ndl;

static int failures = 0;
static void report(const char* label, bool ok)
{
    std::printf("%s: %s\n", label, ok ? "Pass" : "Fail");
    std::fflush(stdout);
    if (!ok) failures++;
}

// image_convolution_tests.cpp:13-67
static void image_convolution()
{
    std::vector<int> data(16);
    Image<int, 2> img(data.data(), {4, 4});
    int i = 0;
    for (auto it = img.begin(); it != img.end(); ++it) *it = ++i;
    std::vector<int> idk = {0, 0, 0, 0, 1, 0, 0, 0, 0};
    Image<int, 2> identity(idk.data(), {3, 3});
    std::vector<int> outData(16);
    Image<int, 2> out(outData.data(), {4, 4});
    convolve(img, out, identity);
    report("convolve() with identity kernel leaves image unchanged", out == img);
    std::vector<int> sk(9, 1);
    Image<int, 2> sumKernel(sk.data(), {3, 3});
    convolve(img, out, sumKernel, BorderMode::Clamp);
    report("convolve() sum kernel matches expected interior value", out(1, 1) == 1 + 2 + 3 + 5 + 6 + 7 + 9 + 10 + 11);
    int clampCorner = out(0, 0);
    report("convolve() clamp border matches hand-derived value",
           clampCorner == 4 * img(0, 0) + 2 * img(1, 0) + 2 * img(0, 1) + img(1, 1));
    convolve(img, out, sumKernel, BorderMode::Wrap);
    report("convolve() wrap border differs from clamp border at the corner", out(0, 0) != clampCorner);
    std::vector<int> sk5(25, 1);
    Image<int, 2> sumKernel5(sk5.data(), {5, 5});
    convolve(img, out, sumKernel5, BorderMode::Clamp);
    int c5 = out(0, 0);
    convolve(img, out, sumKernel5, BorderMode::Reflect);
    report("convolve() reflect border differs from clamp border with a wider kernel", out(0, 0) != c5);
}

// composite_tests.cpp:105-133
static void convolve_then_reduce()
{
    std::vector<int> data(20);
    Image<int, 2> img(data.data(), {5, 4});
    int i = 0;
    for (auto it = img.begin(); it != img.end(); ++it) *it = ++i;
    std::vector<int> kd = {1, 2, 1, 2, 4, 2, 1, 2, 1};
    Image<int, 2> kernel(kd.data(), {3, 3});
    std::vector<int> od(20);
    Image<int, 2> out(od.data(), {5, 4});
    convolve(img, out, kernel, BorderMode::Wrap);
    long so = 0, si = 0, sk = 0;
    for (int v : od) so += v;
    for (int v : data) si += v;
    for (int v : kd) sk += v;
    report("convolve() with wrap border + sum() satisfies kernel.sum()*image.sum() identity", so == sk * si);
}

// type_coverage_tests.cpp:141-149
static void float_average()
{
    std::vector<float> ad = {3, 5, 7, 9};
    Image<float, 1> a(ad.data(), {4});
    OwnedImage<double, 1> avgKernel({3}, {1.0 / 3, 1.0 / 3, 1.0 / 3});
    std::array<float, 4> od{};
    Image<float, 1> out(od.data(), {4});
    convolve(a, out, avgKernel, BorderMode::Clamp);
    report("float convolve() average equals (3+5+7)/3", out(1) == (3.0f + 5.0f + 7.0f) / 3.0f);
}

// convolution_tutorial.md:61-66,83-85,138-149
static void tutorial()
{
    OwnedImage<int, 2> grid({5, 5});
    for (int y = 0; y < 5; y++)
        for (int x = 0; x < 5; x++) grid(x, y) = 5 * y + x + 1;
    OwnedImage<double, 2> sumKernel({3, 3});
    make_box_kernel(sumKernel);
    OwnedImage<int, 2> out2({5, 5});
    convolve(grid, out2, sumKernel);
    const int table[5][5] = {{27, 33, 42, 51, 57}, {57, 63, 72, 81, 87}, {102, 108, 117, 126, 132},
                             {147, 153, 162, 171, 177}, {177, 183, 192, 201, 207}};
    bool ok = true;
    for (int y = 0; y < 5; y++)
        for (int x = 0; x < 5; x++) ok &= out2(x, y) == table[y][x];
    report("tutorial sum-kernel table", ok);
    convolve(grid, out2, sumKernel, BorderMode::Wrap);
    bool w = out2(0, 0) == 99;
    convolve(grid, out2, sumKernel, BorderMode::Reflect);
    report("tutorial corners clamp 27 / wrap 99 / reflect 27", w && out2(0, 0) == 27);
    OwnedImage<double, 2> impulse({11, 11});
    impulse = 0.0;
    impulse(5, 5) = 1000.0;
    OwnedImage<double, 2> response({11, 11});
    gaussian_blur(impulse, response, 1.0);
    report("tutorial gaussian impulse centre 159.24",
           std::round(response(5, 5) * 100) == 15924 && std::round(response(4, 5) * 100) == 9658 &&
               std::round(response(4, 4) * 100) == 5858);
}

// per_channel over interleaved {3,W,H} colour (demo/convolution/convolution.cpp:75-82)
static void per_channel_box()
{
    const int W = 37, H = 23;
    OwnedImage<uint8_t, 3> src({3, W, H});
    std::mt19937 rng(5);
    for (auto it = src.begin(); it != src.end(); ++it) *it = (uint8_t)(rng() & 255);
    OwnedImage<uint8_t, 3> dst({3, W, H});
    OwnedImage<double, 2> box({7, 7});
    make_box_kernel(box);
    per_channel(src, dst, 0, [&](const auto& s, auto& d) { convolve(s, d, box, BorderMode::Wrap); });
    bool ok = true;
    for (int c = 0; c < 3; c++)
        for (int y = 0; y < H; y++)
            for (int x = 0; x < W; x++) {
                int s = 0;
                for (int dy = -3; dy <= 3; dy++)
                    for (int dx = -3; dx <= 3; dx++) s += src(c, (x + dx + W) % W, (y + dy + H) % H);
                ok &= dst(c, x, y) == (uint8_t)(s & 255);
            }
    report("per_channel 7x7 box on interleaved uint8 equals window sum mod 256", ok);
    // the same through two separable passes on the strided channel views
    OwnedImage<uint8_t, 3> dst2({3, W, H});
    per_channel(src, dst2, 0, [&](const auto& s, auto& d) {
        convolve_separable<Image<uint8_t, 2>, Image<uint8_t, 2>, 2>(
            s, d, {std::vector<double>(7, 1.0), std::vector<double>(7, 1.0)}, {BorderMode::Wrap, BorderMode::Wrap});
    });
    report("separable 7-tap box passes equal the dense 7x7 box (uint8, wrap)", dst2 == dst);
}

// image_view_tests.cpp semantics: convolving a view equals convolving its copy
static void views()
{
    OwnedImage<float, 3> base({21, 17, 13});
    std::mt19937 rng(7);
    std::uniform_real_distribution<float> u(0, 1);
    for (auto it = base.begin(); it != base.end(); ++it) *it = u(rng);
    auto views = {base.mirror(0), base.view({1, 2, 0}, {-2, -1, -1}, {2, -1, 1}), base.swap_axes(0, 2)};
    bool ok = true;
    for (const auto& v : views) {
        OwnedImage<float, 3> copy(v);
        OwnedImage<float, 3> a(v.extent()), b(v.extent());
        gaussian_blur(v, a, 0.9, BorderMode::Reflect);
        gaussian_blur(copy, b, 0.9, BorderMode::Reflect);
        ok &= a == b;
    }
    report("gaussian_blur of a strided view equals that of its dense copy", ok);
    auto sl = base.slice(2, 5);
    OwnedImage<float, 2> a(sl.extent()), b(sl.extent());
    OwnedImage<float, 2> copy(sl);
    gaussian_blur(sl, a, 1.3, BorderMode::Wrap);
    gaussian_blur(copy, b, 1.3, BorderMode::Wrap);
    report("gaussian_blur of a slice equals that of its dense copy", a == b);
    // writing through a strided destination view leaves other elements alone
    OwnedImage<float, 3> whole({21, 17, 13});
    whole = -1.0f;
    auto dv = whole.view({0, 0, 0}, {-1, -1, -1}, {1, 2, 1});
    OwnedImage<float, 3> srcv(dv.extent());
    srcv = 2.0f;
    gaussian_blur(srcv, dv, 1.0, BorderMode::Clamp);
    bool untouched = true, written = true;
    for (const auto& c : whole.coordinates()) {
        if (c[1] % 2) untouched &= whole.at(c) == -1.0f;
        else written &= std::fabs(whole.at(c) - 2.0f) < 1e-6f;
    }
    report("strided destination: only the view's elements are written", untouched && written);
}

// per-axis sigma and the FP32 contract (1e-6 relative on U[0,1) data)
static void per_axis_fp32()
{
    OwnedImage<float, 3> x({40, 33, 29});
    std::mt19937 rng(9);
    std::uniform_real_distribution<float> u(0, 1);
    for (auto it = x.begin(); it != x.end(); ++it) *it = u(rng);
    OwnedImage<float, 3> ref(x.extent()), f32(x.extent());
    std::array<double, 3> sig = {1.5, 1.5, 3.0};
    std::array<BorderMode, 3> br = {BorderMode::Reflect, BorderMode::Reflect, BorderMode::Reflect};
    gaussian_blur<OwnedImage<float, 3>, OwnedImage<float, 3>, 3>(x, ref, sig, br, Accumulation::Reference);
    gaussian_blur<OwnedImage<float, 3>, OwnedImage<float, 3>, 3>(x, f32, sig, br, Accumulation::FP32);
    double worst = 0;
    for (const auto& c : x.coordinates()) worst = std::max(worst, std::fabs((double)f32.at(c) - ref.at(c)) / ref.at(c));
    report("per-axis Gaussian: FP32 accumulation within 1e-6 of the reference arithmetic", worst <= 1e-6);
    DeviceImage<float, 3> dx(x), dy(x.extent());
    gaussian_blur<DeviceImage<float, 3>, DeviceImage<float, 3>, 3>(dx, dy, sig, br, Accumulation::FP32);
    report("DeviceImage path equals the staged host path", dy.to_host() == f32);
}

// an image type of the caller's own (not this library's Image) with the
// reference's minimal interface plus strides (image/detail.h:48-57,90-98):
// one channel of an interleaved RGB buffer
template <class T> struct ChannelView {
    typedef T value_type;
    T* base;
    int w, h, ch;
    std::array<int, 2> extent() const { return {w, h}; }
    std::array<int, 2> stride() const { return {3, 3 * w}; }
    T& at(const std::array<int, 2>& i) const { return base[ch + 3 * (i[0] + w * i[1])]; }
    std::vector<std::array<int, 2>> coordinates() const
    {
        std::vector<std::array<int, 2>> c;
        for (int y = 0; y < h; y++)
            for (int x = 0; x < w; x++) c.push_back({x, y});
        return c;
    }
};

static void foreign_image_type()
{
    const int w = 53, h = 41;
    std::vector<float> rgb(3 * w * h), out(3 * w * h, -7.f);
    std::mt19937 rng(13);
    std::uniform_real_distribution<float> u(0, 1);
    for (float& v : rgb) v = u(rng);
    ChannelView<float> src{rgb.data(), w, h, 1}, dst{out.data(), w, h, 1};
    gaussian_blur(src, dst, 1.5, BorderMode::Reflect);
    // the same channel through this library's strided Image view
    Image<float, 3> all(rgb.data(), {3, w, h});
    OwnedImage<float, 2> want({w, h});
    gaussian_blur(all.slice(0, 1), want, 1.5, BorderMode::Reflect);
    bool ok = true;
    for (const auto& c : want.coordinates()) ok = ok && dst.at(c) == want.at(c);
    for (int i = 0; i < w * h; i++) ok = ok && out[3 * i] == -7.f && out[3 * i + 2] == -7.f;
    report("a caller's own strided image type (minimal interface + stride access) works unchanged", ok);
}

// NRRD / raw files straight into device images through pinned host memory
static void file_staging()
{
    OwnedImage<float, 3> x({37, 30, 21});
    std::mt19937 rng(11);
    std::uniform_real_distribution<float> u(0, 1);
    for (auto it = x.begin(); it != x.end(); ++it) *it = u(rng);
    const std::string nrrd = "/tmp/ndl_dropin_volume.nrrd", raw = "/tmp/ndl_dropin_volume.raw";
    image_io::save(x, nrrd);
    image_io::save_raw(x, raw);
    DeviceImage<float, 3> dn = image_io::load_nrrd_device<float, 3>(nrrd);
    DeviceImage<float, 3> dr = image_io::load_raw_device<float, 3>(raw, x.extent(), 0);
    report("load_nrrd_device / load_raw_device upload the file's values", dn.to_host() == x && dr.to_host() == x);
    std::array<double, 3> sig = {1.0, 2.0, 1.5};
    std::array<BorderMode, 3> br = {BorderMode::Clamp, BorderMode::Reflect, BorderMode::Wrap};
    OwnedImage<float, 3> host(x.extent());
    gaussian_blur<OwnedImage<float, 3>, OwnedImage<float, 3>, 3>(x, host, sig, br, Accumulation::FP32);
    DeviceImage<float, 3> dy(x.extent());
    gaussian_blur<DeviceImage<float, 3>, DeviceImage<float, 3>, 3>(dn, dy, sig, br, Accumulation::FP32);
    image_io::save(dy, "/tmp/ndl_dropin_blurred.nrrd");
    std::array<int, 3> e{};
    std::vector<float> back = image_io::load_nrrd<float, 3>(e, "/tmp/ndl_dropin_blurred.nrrd");
    report("blur of a loaded device volume, saved from the device, equals the host path",
           e == x.extent() && Image<float, 3>(back.data(), e) == host);
    // a uint8 file converted to float on the way in
    OwnedImage<unsigned char, 2> g({64, 48});
    int i = 0;
    for (auto it = g.begin(); it != g.end(); ++it) *it = (unsigned char)(i++ * 13);
    image_io::save(g, "/tmp/ndl_dropin_u8.nrrd");
    DeviceImage<float, 2> dg = image_io::load_nrrd_device<float, 2>("/tmp/ndl_dropin_u8.nrrd");
    OwnedImage<float, 2> hg = dg.to_host();
    bool ok = true;
    for (const auto& c : g.coordinates()) ok = ok && hg.at(c) == (float)g.at(c);
    report("load_nrrd_device converts a uint8 file to float", ok);
}

// adjoints: <A x, w> == <x, A^T w>, and tap gradients vs central differences
static void adjoints()
{
    const std::array<int, 3> e = {19, 14, 11};
    OwnedImage<double, 3> x(e), w(e), ax(e), atw(e);
    std::mt19937 rng(11);
    std::normal_distribution<double> n(0, 1);
    for (auto it = x.begin(); it != x.end(); ++it) *it = n(rng);
    for (auto it = w.begin(); it != w.end(); ++it) *it = n(rng);
    std::array<std::vector<double>, 3> taps;
    for (int a = 0; a < 3; a++)
        for (int t = 0; t < 5 + 2 * a; t++) taps[a].push_back(n(rng));
    for (BorderMode b : {BorderMode::Clamp, BorderMode::Wrap, BorderMode::Reflect}) {
        std::array<BorderMode, 3> br = {b, b, b};
        convolve_separable<OwnedImage<double, 3>, OwnedImage<double, 3>, 3>(x, ax, taps, br);
        convolve_adjoint<OwnedImage<double, 3>, OwnedImage<double, 3>, 3>(w, atw, taps, br);
        double lhs = 0, rhs = 0;
        for (const auto& c : x.coordinates()) {
            lhs += ax.at(c) * w.at(c);
            rhs += x.at(c) * atw.at(c);
        }
        report("adjoint dot-product identity <Ax,w> == <x,A^T w>", std::fabs(lhs - rhs) <= 1e-10 * std::fabs(lhs));
        auto g = convolve_adjoint_taps<OwnedImage<double, 3>, OwnedImage<double, 3>, 3>(x, w, taps, br);
        bool ok = true;
        for (int a = 0; a < 3; a++)
            for (std::size_t t = 0; t < taps[a].size(); t++) {
                auto tp = taps, tm = taps;
                tp[a][t] += 1e-3;
                tm[a][t] -= 1e-3;
                OwnedImage<double, 3> yp(e), ym(e);
                convolve_separable<OwnedImage<double, 3>, OwnedImage<double, 3>, 3>(x, yp, tp, br);
                convolve_separable<OwnedImage<double, 3>, OwnedImage<double, 3>, 3>(x, ym, tm, br);
                double lp = 0, lm = 0;
                for (const auto& c : x.coordinates()) {
                    lp += w.at(c) * yp.at(c);
                    lm += w.at(c) * ym.at(c);
                }
                double fd = (lp - lm) / 2e-3;
                ok &= std::fabs(fd - g[a][t]) <= 1e-7 * std::max(1.0, std::fabs(fd));
            }
        report("tap gradients equal FP64 central differences", ok);
    }
}

// image_convolution_tests.cpp:117-160: ramp -> 1 in the interior; plane 2x+3y -> (2,3)
static void gradients()
{
    OwnedImage<double, 2> plane({9, 7});
    for (int y = 0; y < 7; y++)
        for (int x = 0; x < 9; x++) plane(x, y) = 2.0 * x + 3.0 * y;
    OwnedImage<double, 3> g({2, 9, 7});
    gradient(plane, g);
    bool ok = true;
    for (int y = 1; y < 6; y++)
        for (int x = 1; x < 8; x++) ok &= std::fabs(g(0, x, y) - 2.0) < 1e-9 && std::fabs(g(1, x, y) - 3.0) < 1e-9;
    report("gradient() of the plane 2x+3y is (2,3) in the interior", ok);
    // Reflect at the edge: x=0 sees x=-1 -> 0, so d/dx = 0.5*(x1 - x0)
    report("gradient() Reflect border at x=0", std::fabs(g(0, 0, 3) - 0.5 * (plane(1, 3) - plane(0, 3))) < 1e-12);
}

// image_morphology_tests.cpp:14-98
static void image_morphology()
{
    std::vector<int> gridData(25);
    Image<int, 2> grid(gridData.data(), {5, 5});
    int i = 0;
    for (auto it = grid.begin(); it != grid.end(); ++it) *it = ++i;
    std::vector<double> boxData(9), crossData(9);
    Image<double, 2> box(boxData.data(), {3, 3}), cross(crossData.data(), {3, 3});
    make_box_kernel(box);
    make_cross_kernel(cross);
    std::vector<int> outData(25);
    Image<int, 2> out(outData.data(), {5, 5});
    erode(grid, out, box);
    report("erode(box) takes the minimum of the 3x3 neighborhood", out(2, 2) == 7);
    dilate(grid, out, box);
    report("dilate(box) takes the maximum of the 3x3 neighborhood", out(2, 2) == 19);
    erode(grid, out, cross);
    report("erode(cross) considers only the 5-tap plus-sign neighborhood", out(2, 2) == 8);
    dilate(grid, out, cross);
    report("dilate(cross) considers only the 5-tap plus-sign neighborhood", out(2, 2) == 18);

    std::vector<int> dotData(25, 0), dotOutData(25);
    Image<int, 2> dot(dotData.data(), {5, 5}), dotOut(dotOutData.data(), {5, 5});
    dot(2, 2) = 1;
    dilate(dot, dotOut, box);
    int ones = 0;
    for (auto it = dotOut.begin(); it != dotOut.end(); ++it) ones += *it == 1;
    report("dilate(box) of a single point turns on all 9 taps of the box", ones == 9);
    dilate(dot, dotOut, cross);
    ones = 0;
    for (auto it = dotOut.begin(); it != dotOut.end(); ++it) ones += *it == 1;
    report("dilate(cross) of a single point turns on only the 5 cross taps", ones == 5);

    std::vector<int> stageData(25), manualData(25), openData(25), interData(25);
    Image<int, 2> stage(stageData.data(), {5, 5}), manual(manualData.data(), {5, 5});
    Image<int, 2> opened(openData.data(), {5, 5}), inter(interData.data(), {5, 5});
    erode(grid, stage, box);
    dilate(stage, manual, box);
    morphological_open(grid, opened, inter, box);
    report("morphological_open() matches erode() then dilate() done by hand", opened == manual);
    dilate(grid, stage, box);
    erode(stage, manual, box);
    morphological_close(grid, opened, inter, box);
    report("morphological_close() matches dilate() then erode() done by hand", opened == manual);

    // 3-D box on a strided view equals the same call on its dense copy
    std::mt19937 rng(7);
    std::vector<float> vol(12 * 10 * 8);
    for (auto& v : vol) v = std::uniform_real_distribution<float>(-1, 1)(rng);
    Image<float, 3> V(vol.data(), {12, 10, 8});
    auto view = V.view({1, 0, 1}, {11, 9, 7}, {2, 1, 1});
    OwnedImage<float, 3> dense(view.extent()), a(view.extent()), b(view.extent());
    dense.assign(view);
    std::vector<double> k3(3 * 3 * 5, 1.0);
    Image<double, 3> box3(k3.data(), {3, 3, 5});
    dilate(view, a, box3, BorderMode::Reflect);
    dilate(dense, b, box3, BorderMode::Reflect);
    report("dilate() of a strided view equals dilate() of its dense copy", a == b);
}

// image_convolution_tests.cpp:70-91
static void downsample_tests()
{
    OwnedImage<uint8_t, 3> constColor({3, 40, 30});
    constColor = uint8_t(77);
    auto constDown = downsample(constColor, 2);
    report("downsample() extent: channel axis untouched, spatial axes halved",
           constDown.extent() == std::array<int, 3>{3, 20, 15});
    bool constMatches = true;
    for (const auto& c : constDown.coordinates())
        if (constDown.at(c) != 77) constMatches = false;
    report("downsample() of a constant image stays constant", constMatches);
    auto explicitAxis = downsample(constColor, 2, 0, BorderMode::Clamp);
    report("downsample() default channelAxis matches explicit channelAxis=0", explicitAxis == constDown);
    // a ramp: 2x2 block means of 0..(w*h-1) are exact halves
    OwnedImage<float, 2> ramp({8, 6});
    int i = 0;
    for (auto it = ramp.begin(); it != ramp.end(); ++it) *it = (float)i++;
    auto half = downsample(ramp, 2, 1);
    bool ok = half.extent() == std::array<int, 2>{4, 6};
    for (int y = 0; y < 6 && ok; y++)
        for (int x = 0; x < 4; x++) ok = ok && half(x, y) == (float)(y * 8 + 2 * x) + 0.5f;
    report("downsample() along x only (channelAxis 1) averages pairs", ok);
    // one pyramid step in one call equals gaussian_blur() then downsample()
    std::mt19937 rng(12);
    OwnedImage<uint8_t, 3> colour({3, 45, 31});
    for (auto it = colour.begin(); it != colour.end(); ++it) *it = (uint8_t)(rng() & 255);
    OwnedImage<uint8_t, 3> blurred(colour.extent());
    gaussian_blur(colour, blurred, 1.0);
    auto twoCalls = downsample(blurred, 2);
    auto oneCall = gaussian_downsample(colour, 1.0, 2);
    report("gaussian_downsample() equals gaussian_blur() then downsample() (uint8 colour)", oneCall == twoCalls);
    OwnedImage<float, 2> gray({37, 29});
    for (auto it = gray.begin(); it != gray.end(); ++it) *it = std::uniform_real_distribution<float>(0, 1)(rng);
    OwnedImage<float, 2> gb(gray.extent());
    gaussian_blur(gray, gb, 1.7, BorderMode::Reflect);
    report("gaussian_downsample() equals gaussian_blur() then downsample() (float, factor 3, reflect)",
           gaussian_downsample(gray, 1.7, 3, 1, BorderMode::Reflect) == downsample(gb, 3, 1));
}

// the scale stack and DoG layers of detect_keypoints (feature_detection.h:91-120)
static void scale_stack()
{
    std::mt19937 rng(11);
    OwnedImage<double, 2> src({70, 50});
    for (auto it = src.begin(); it != src.end(); ++it) *it = std::uniform_real_distribution<double>(0, 255)(rng);
    const int numScales = 5;
    const double sigma0 = 1.6;
    std::vector<double> sig;
    std::vector<OwnedImage<double, 2>> levels, singles, dogs;
    for (int i = 0; i < numScales; i++) {
        sig.push_back(sigma0 * std::pow(2.0, (double)i / (numScales - 3)));
        levels.emplace_back(src.extent());
        singles.emplace_back(src.extent());
        gaussian_blur(src, singles.back(), sig.back(), BorderMode::Reflect);
    }
    gaussian_blur_scales(src, levels, sig, BorderMode::Reflect);
    bool same = true;
    for (int i = 0; i < numScales; i++) same = same && levels[i] == singles[i];
    report("gaussian_blur_scales() equals one gaussian_blur() per sigma", same);
    for (int i = 0; i + 1 < numScales; i++) dogs.emplace_back(src.extent());
    difference_of_gaussians(levels, dogs);
    bool ok = true;
    for (int i = 0; i + 1 < numScales; i++)
        for (const auto& c : dogs[i].coordinates()) ok = ok && dogs[i].at(c) == levels[i + 1].at(c) - levels[i].at(c);
    report("difference_of_gaussians() layers are level[i+1] - level[i] in double", ok);
    // integral images round like gaussian_blur (std::round, DIM >= 2)
    OwnedImage<uint8_t, 2> u8({64, 48});
    for (auto it = u8.begin(); it != u8.end(); ++it) *it = (uint8_t)std::uniform_int_distribution<int>(0, 255)(rng);
    std::vector<OwnedImage<uint8_t, 2>> ul, us;
    for (int i = 0; i < 3; i++) {
        ul.emplace_back(u8.extent());
        us.emplace_back(u8.extent());
        gaussian_blur(u8, us.back(), sig[i], BorderMode::Clamp);
    }
    gaussian_blur_scales(u8, ul, std::vector<double>(sig.begin(), sig.begin() + 3), BorderMode::Clamp);
    same = true;
    for (int i = 0; i < 3; i++) same = same && ul[i] == us[i];
    report("gaussian_blur_scales() on uint8 equals gaussian_blur() per sigma", same);
}

int main()
{
    int n = 0;
    if (ndl_device_count(&n) != 0 || n < 1) {
        std::printf("no GPU visible\n");
        return 99;
    }
    image_convolution();
    convolve_then_reduce();
    float_average();
    tutorial();
    per_channel_box();
    views();
    per_axis_fp32();
    adjoints();
    gradients();
    image_morphology();
    downsample_tests();
    scale_stack();
    file_staging();
    foreign_image_type();
    std::printf("%d failures\n", failures);
    return failures;
}
