// Self-contained test runner of the C++ API (include/dt/), mirroring the
// reference's gtest suites (gtest is not available here):
//   CPU : tests/image_view.cpp, tests/image2d.cpp, tests/c4.cpp, tests/rgb8.cpp,
//         tests/io.cpp, plus normalize / inferno / invert_mask
//   GPU : tests/transfert.cpp, tests/distance_transform.cpp (the two
//         FastGeodis golden vectors, bit-exact, and the raster == blocks ==
//         task cross-schedule checks), plus the extension overloads.
// Images compare with exact per-pixel equality (tests/helpers.hpp:10-37).
// The reference draws its random test images from std::random_device; this
// runner uses a seeded SplitMix64 so a failure reproduces.
//
// Usage: dt_tests [--gpu] [--golden DIR]
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <fstream>
#include <functional>
#include <sstream>
#include <string>
#include <vector>

#include <dt/block_priority.hpp>
#include <dt/fill.hpp>
#include <dt/generalised_distance_transform.hpp>
#include <dt/image2d.hpp>
#include <dt/imread.hpp>
#include <dt/imsave.hpp>
#include <dt/inferno.hpp>
#include <dt/invert_mask.hpp>
#include <dt/normalize.hpp>
#include <dt/neighborhood.hpp>
#include <dt/rgb.hpp>
#include <dt/transfert.hpp>
#include <dtk.h>

namespace
{
  int         g_fail = 0, g_run = 0;
  std::string g_golden = "tests/golden";

  struct TestFailure
  {
    std::string msg;
  };

#define EXPECT(cond)                                                                                                  \
  do                                                                                                                  \
  {                                                                                                                   \
    if (!(cond))                                                                                                      \
      throw TestFailure{std::string(__FILE__) + ":" + std::to_string(__LINE__) + ": " #cond};                         \
  } while (0)

  void run_test(const char* name, const std::function<void()>& f)
  {
    ++g_run;
    std::printf("[ RUN      ] %s\n", name);
    std::fflush(stdout);
    try
    {
      f();
      std::printf("[       OK ] %s\n", name);
    }
    catch (const TestFailure& e)
    {
      ++g_fail;
      std::printf("[  FAILED  ] %s\n  %s\n", name, e.msg.c_str());
    }
    catch (const std::exception& e)
    {
      ++g_fail;
      std::printf("[  FAILED  ] %s\n  exception: %s\n", name, e.what());
    }
    std::fflush(stdout);
  }

  struct SplitMix64
  {
    std::uint64_t s;
    std::uint64_t next()
    {
      std::uint64_t z = (s += 0x9E3779B97F4A7C15ull);
      z               = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
      z               = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
      return z ^ (z >> 31);
    }
  };

  template <typename T>
  dt::image2d<T> random_image(int w, int h, std::uint64_t seed)
  {
    SplitMix64     r{seed};
    dt::image2d<T> img(w, h);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x)
      {
        std::uint64_t v = r.next();
        std::memcpy(&img(x, y), &v, sizeof(T));
      }
    return img;
  }

  template <typename T, typename V>
  void expect_images_eq(const dt::image2d_view<T>& a, const dt::image2d_view<V>& b)
  {
    EXPECT(a.width() == b.width() && a.height() == b.height());
    EXPECT(a.memory_kind() == dt::e_memory_kind::CPU && b.memory_kind() == dt::e_memory_kind::CPU);
    for (int x = 0; x < a.width(); x++)
      for (int y = 0; y < a.height(); y++)
        if (!(a(x, y) == b(x, y)))
        {
          std::ostringstream o;
          o << "img1(" << x << ", " << y << ") (" << +a(x, y) << ") != img2(" << x << ", " << y << ") (" << +b(x, y)
            << ")";
          throw TestFailure{o.str()};
        }
  }

  // Minimal reader of the golden fixture JSON (flat numeric arrays / scalars).
  struct Fixture
  {
    std::string text;
    explicit Fixture(const std::string& path)
    {
      std::ifstream f(path);
      if (!f)
        throw std::runtime_error("cannot open fixture " + path);
      std::stringstream ss;
      ss << f.rdbuf();
      text = ss.str();
    }
    std::vector<double> array(const std::string& key) const
    {
      const auto k = text.find("\"" + key + "\"");
      if (k == std::string::npos)
        throw std::runtime_error("fixture key " + key);
      const auto          a = text.find('[', k), b = text.find(']', a);
      std::vector<double> out;
      std::stringstream   ss(text.substr(a + 1, b - a - 1));
      std::string         tok;
      while (std::getline(ss, tok, ','))
        out.push_back(std::stod(tok));
      return out;
    }
    double scalar(const std::string& key) const
    {
      const auto k = text.find("\"" + key + "\"");
      const auto c = text.find(':', k);
      return std::stod(text.substr(c + 1));
    }
  };

  dt::image2d<float> from_bits(const std::vector<double>& bits, int w, int h)
  {
    dt::image2d<float> r(w, h);
    for (int y = 0; y < h; ++y)
      for (int x = 0; x < w; ++x)
      {
        const std::uint32_t u = (std::uint32_t)bits[(size_t)y * w + x];
        float               f;
        std::memcpy(&f, &u, 4);
        r(x, y) = f;
      }
    return r;
  }

  void cpu_tests()
  {
    run_test("Image2DView.default_constructor", [] {
      dt::image2d_view<void> img;
      EXPECT(img.width() == 0 && img.height() == 0 && img.pitch() == 0 && img.buffer() == nullptr && !img.valid());
    });
    run_test("Image2DView.argument_constructor_and_assignment_operators", [] {
      std::uint8_t           data[6];
      dt::image2d_view<void> img(data, 3, 2, 3, 1);
      EXPECT(img.width() == 3 && img.height() == 2 && img.pitch() == 3 && img.elem_size() == 1 && img.valid());
      auto img2 = dt::image2d_view<void>(img);
      EXPECT(img2.width() == 3 && img2.buffer() != nullptr);
      auto img3 = dt::image2d_view<void>(std::move(img2));
      EXPECT(img2.buffer() == nullptr && !img2.valid());
      EXPECT(img3.width() == 3 && img3.height() == 2 && img3.pitch() == 3 && img3.valid());
      auto img4 = img;
      EXPECT(img4.valid() && img4.pitch() == 3);
      auto img5 = std::move(img4);
      EXPECT(img4.buffer() == nullptr && img5.valid());
      EXPECT(!img.in_domain(-1, 0) && !img.in_domain(0, -1) && !img.in_domain(-1, -1) && !img.in_domain(3, 2) &&
             img.in_domain(0, 0));
    });
    run_test("Image2DView.TypedView", [] {
      std::uint16_t    data[] = {5, 9, 7, 3, 4, 6};
      dt::image2d_view img(data, 3, 2, 6);
      EXPECT(img.width() == 3 && img.height() == 2 && img.pitch() == 6 && img.elem_size() == 2 && img.valid());
      EXPECT(img(0, 0) == 5 && img(1, 0) == 9 && img(2, 0) == 7 && img(0, 1) == 3 && img(1, 1) == 4 && img(2, 1) == 6);
    });
    run_test("Image2D.uint8_cpu", [] {
      std::uint8_t              data[] = {3, 8, 5, 2, 1, 9};
      dt::image2d<std::uint8_t> img(3, 2);
      EXPECT(img.width() == 3 && img.height() == 2 && img.pitch() == 3 && img.elem_size() == 1 && img.valid());
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 3; x++)
          img(x, y) = data[y * 3 + x];
      for (int y = 0; y < 2; y++)
        for (int x = 0; x < 3; x++)
          EXPECT(img(x, y) == data[y * 3 + x]);
    });
    run_test("Image2D.uint16_cpu", [] {
      dt::image2d<std::uint16_t> img(3, 2);
      EXPECT(img.pitch() == 6 && img.elem_size() == 2);
      auto copy = img; // shares storage
      img(1, 1) = 42;
      EXPECT(copy(1, 1) == 42);
    });
    run_test("C4.c4", [] {
      const dt::point2d ref[] = {{2, 7}, {3, 6}, {4, 7}, {3, 8}};
      int               i     = 0;
      for (const auto& n : dt::c4(dt::point2d(3, 7)))
        EXPECT(n == ref[i++]);
      EXPECT(i == 4);
    });
    run_test("C8.c8", [] {
      const dt::point2d ref[] = {{-1, 0}, {-1, -1}, {0, -1}, {1, -1}, {1, 0}, {1, 1}, {0, 1}, {-1, 1}};
      int               i     = 0;
      for (const auto& n : dt::c8(dt::point2d(0, 0)))
        EXPECT(n == ref[i++]);
      EXPECT(i == 8);
    });
    run_test("Color.Rgb8", [] {
      dt::rgb8 c{1, 4, 3};
      EXPECT(c.r == 1 && c.g == 4 && c.b == 3 && sizeof(dt::rgb8) == 3);
    });
    run_test("Io.uint8", [] { // tests/io.cpp:8-16 of the reference
      const std::uint8_t data[] = {10, 3, 5, 8, 1, 3, 0, 4, 10, 4};
      dt::image2d_view   in(data, 5, 2, 5);
      dt::imsave("test_tmp_io.png", in);
      auto out = dt::imread<std::uint8_t>("test_tmp_io.png");
      expect_images_eq(in, out);
      std::remove("test_tmp_io.png");
    });
    run_test("Io.pitched_uint8", [] {
      auto img = random_image<std::uint8_t>(37, 11, 5);
      // a view with padded rows: every other column of a wider buffer's rows is skipped by the pitch
      std::vector<std::uint8_t> wide(64 * 11, 0);
      for (int y = 0; y < 11; ++y)
        for (int x = 0; x < 37; ++x)
          wide[y * 64 + x] = img(x, y);
      dt::image2d_view<std::uint8_t> view(wide.data(), 37, 11, 64);
      dt::imsave("test_tmp_io2.png", view);
      expect_images_eq(img, dt::imread<std::uint8_t>("test_tmp_io2.png"));
      std::remove("test_tmp_io2.png");
    });
    run_test("Io.errors", [] {
      bool t1 = false, t2 = false, t3 = false;
      const std::uint8_t d[4] = {1, 2, 3, 4};
      dt::image2d_view   v(d, 2, 2, 2);
      try { dt::imsave("test_tmp_io.jpg", v); } catch (const std::invalid_argument&) { t1 = true; }
      try { dt::imread<std::uint8_t>("/nonexistent/file.png"); } catch (const std::runtime_error&) { t2 = true; }
      dt::image2d<dt::rgb8> c(3, 2);
      dt::fill(c, dt::rgb8{1, 2, 3});
      dt::imsave("test_tmp_rgb.png", c);
      try { dt::imread<std::uint8_t>("test_tmp_rgb.png"); } catch (const std::invalid_argument&) { t3 = true; }
      std::remove("test_tmp_rgb.png");
      EXPECT(t1 && t2 && t3);
    });
    run_test("Normalize.known_answers", [] {
      // [0, 1, 2] -> 0, 1/2 * 255 = 127.5 -> 127, 255; negative values; the
      // reference's max start (smallest positive float) on an all-zero image.
      float        a[] = {0.f, 1.f, 2.f, -1.f, 0.f, 1.f, 0.f, 0.f, 0.f};
      std::uint8_t w[] = {0, 127, 255, 0, 127, 255, 0, 0, 0};
      for (int k = 0; k < 3; ++k)
      {
        dt::image2d_view<float> in(a + 3 * k, 3, 1, 12);
        auto                    out = dt::normalize<std::uint8_t>(in);
        for (int x = 0; x < 3; ++x)
          EXPECT(out(x, 0) == w[3 * k + x]);
      }
      float                   z[] = {0.f, 0.f};
      dt::image2d_view<float> zv(z, 2, 1, 8);
      EXPECT(dt::minmax(zv).first == 0.f && dt::minmax(zv).second > 0.f && dt::minmax(zv).second < 1e-37f);
    });
    run_test("Inferno.table_ends", [] {
      const dt::rgb8* t = dt::inferno_table();
      EXPECT(t[0].r == 0 && t[0].g == 0 && t[0].b == 3 && t[255].r == 252 && t[255].g == 254 && t[255].b == 164);
      std::uint8_t            v[] = {0, 255};
      dt::image2d_view        in(v, 2, 1, 2);
      auto                    out = dt::inferno(in);
      EXPECT(out(0, 0).b == 3 && out(1, 0).r == 252);
    });
    run_test("InvertMask.invert", [] {
      std::uint8_t     m[] = {0, 1, 7, 0};
      dt::image2d_view in(m, 4, 1, 4);
      auto             out = dt::invert_mask(in);
      EXPECT(out(0, 0) == 1 && out(1, 0) == 0 && out(2, 0) == 0 && out(3, 0) == 1);
    });
    run_test("Abi.params_default", [] {
      dtk_params p;
      dtk_params_default(&p);
      EXPECT(p.v == 1e10f && p.connectivity == DTK_C8 && p.lambda == 0.f);
      EXPECT(dtk_tile_count(4096, 4096, 16) == 67 * 293);
    });
  }

  dt::image2d<float> run_sched(const dt::image2d<std::uint8_t>& img, const dt::image2d<std::uint8_t>& mask, float lam,
                               int sched)
  {
    auto d_img  = dt::host_to_device(img);
    auto d_mask = dt::host_to_device(mask);
    if (sched == 0)
      return dt::device_to_host(dt::generalised_distance_transform(d_img, d_mask, lam));
    if (sched == 1)
      return dt::device_to_host(dt::generalised_distance_transform_blocks(d_img, d_mask, lam));
    return dt::device_to_host(dt::generalised_distance_transform_task(d_img, d_mask, lam));
  }

  void gpu_tests()
  {
    run_test("Transfert.id_uint8", [] {
      const auto img    = random_image<std::uint8_t>(20, 20, 1);
      const auto result = dt::device_to_host(dt::host_to_device(img));
      expect_images_eq(img, result);
    });
    run_test("Transfert.id_uint32", [] {
      const auto img    = random_image<std::uint32_t>(20, 20, 2);
      const auto result = dt::device_to_host(dt::host_to_device(img));
      expect_images_eq(img, result);
    });
    for (int sched = 0; sched < 3; ++sched)
    {
      const std::string sfx = sched == 0 ? "" : (sched == 1 ? "Blocks" : "Task");
      run_test(("DistanceTransform.EuclidianFastGeodis" + sfx).c_str(), [sched] {
        Fixture                   fx(g_golden + "/fastgeodis_euclidean.json");
        const int                 W = 50, H = 40;
        dt::image2d<std::uint8_t> img(W, H), mask(W, H);
        dt::fill(img, std::uint8_t(0));
        dt::fill(mask, std::uint8_t(0));
        for (int y = 20; y < 30; y++)
          for (int x = 20; x < 43; x++)
            mask(x, y) = 1;
        const auto ref = from_bits(fx.array("expected_f32_bits"), W, H);
        expect_images_eq(run_sched(img, mask, 0.0f, sched), ref);
      });
      run_test(("DistanceTransform.GeodesicFastGeodis" + sfx).c_str(), [sched] {
        Fixture                   fx(g_golden + "/fastgeodis_geodesic.json");
        const int                 W = 36, H = 34;
        const auto                pix = fx.array("image_u8");
        dt::image2d<std::uint8_t> img(W, H), mask(W, H);
        for (int y = 0; y < H; y++)
          for (int x = 0; x < W; x++)
            img(x, y) = (std::uint8_t)pix[(size_t)y * W + x];
        dt::fill(mask, std::uint8_t(0));
        mask(W / 2, H / 2) = 1;
        const auto ref     = from_bits(fx.array("expected_f32_bits"), W, H);
        expect_images_eq(run_sched(img, mask, 1.0f, sched), ref);
      });
    }
    // tests/distance_transform.cpp:482-559: blocks and task equal raster.
    for (float lam : {0.0f, 1.0f, 0.5f})
      for (int sched = 1; sched < 3; ++sched)
      {
        const std::string name = std::string("DistanceTransform.") + (lam == 0.f ? "Euclidian" : lam == 1.f ? "Geodesic" : "Mixed") +
                                 (sched == 1 ? "Blocks" : "Task");
        run_test(name.c_str(), [lam, sched] {
          const auto                img = random_image<std::uint8_t>(250, 230, 7);
          dt::image2d<std::uint8_t> mask(250, 230);
          dt::fill(mask, std::uint8_t(0));
          mask(125, 115) = 1;
          expect_images_eq(run_sched(img, mask, lam, 0), run_sched(img, mask, lam, sched));
        });
      }
    run_test("Extensions.Rgb8SoftMaskC4", [] {
      const int          W = 130, H = 70;
      dt::image2d<dt::rgb8> img(W, H);
      dt::image2d<float> mask(W, H);
      SplitMix64         r{11};
      for (int y = 0; y < H; y++)
        for (int x = 0; x < W; x++)
        {
          const std::uint64_t v = r.next();
          img(x, y)             = dt::rgb8{(std::uint8_t)v, (std::uint8_t)(v >> 8), (std::uint8_t)(v >> 16)};
          mask(x, y)            = 0.f;
        }
      mask(10, 10) = 1.f;
      mask(100, 50) = 0.5f;
      auto         di = dt::host_to_device(img);
      auto         dm = dt::host_to_device(mask);
      dt::ggdt_options o;
      o.connectivity = 4;
      o.schedule     = dt::e_schedule::raster;
      const auto a   = dt::device_to_host(dt::generalised_distance_transform(di, dm, 0.5f, 1e3f, o));
      o.schedule     = dt::e_schedule::task;
      const auto b   = dt::device_to_host(dt::generalised_distance_transform(di, dm, 0.5f, 1e3f, o));
      expect_images_eq(a, b);
      EXPECT(a(10, 10) == 0.f && a(100, 50) <= 500.f);
    });
    run_test("Display.device_equals_host", [] {
      // normalize + inferno of a distance map on the device equal the host
      // versions (the reference computes them on the host).
      auto img  = random_image<std::uint8_t>(173, 91, 31);
      auto mask = dt::image2d<std::uint8_t>(173, 91);
      dt::fill(mask, std::uint8_t(0));
      mask(17, 40) = 1;
      mask(150, 3) = 1;
      const auto dist   = run_sched(img, mask, 0.5f, 2);
      const auto d_dist = dt::host_to_device(dist);
      const auto norm   = dt::normalize<std::uint8_t>(dist);
      const auto d_norm = dt::normalize<std::uint8_t>(d_dist);
      expect_images_eq(norm, dt::device_to_host(d_norm));
      const auto col   = dt::inferno(norm);
      const auto d_col = dt::device_to_host(dt::inferno(d_norm));
      EXPECT(col.width() == d_col.width() && col.height() == d_col.height());
      for (int y = 0; y < col.height(); ++y)
        for (int x = 0; x < col.width(); ++x)
          EXPECT(col(x, y).r == d_col(x, y).r && col(x, y).g == d_col(x, y).g && col(x, y).b == d_col(x, y).b);
      const auto mm = dt::minmax(dist), d_mm = dt::minmax(d_dist);
      EXPECT(mm.first == d_mm.first && mm.second == d_mm.second);
    });
    run_test("BlockPriority.single_seed", [] {
      // One seed pixel in tile (1, 1) of a 3 x 4 grid of 32 x 32 tiles:
      // city-block distances 0..3 hold 1, 4, 5 and 2 tiles, so cdf is 1, 5,
      // 10, 12 and the priorities cdf * 64 / 13 are 4, 24, 49, 59
      // (dt/src/block_priority.cu:109-121).
      dt::image2d<std::uint8_t> mask(70, 100);
      dt::fill(mask, std::uint8_t(0));
      mask(50, 40) = 1;
      const auto     dm = dt::host_to_device(mask);
      std::uint8_t*  prio = nullptr;
      std::uint32_t* cdf  = nullptr;
      EXPECT(hipMalloc(&prio, 12) == hipSuccess && hipMalloc(&cdf, 64 * sizeof(std::uint32_t)) == hipSuccess);
      const std::uint32_t level0 = dt::compute_block_priorities(dm, prio, cdf);
      std::uint8_t        hp[12];
      std::uint32_t       hc[64];
      EXPECT(hipMemcpy(hp, prio, 12, hipMemcpyDeviceToHost) == hipSuccess);
      EXPECT(hipMemcpy(hc, cdf, sizeof(hc), hipMemcpyDeviceToHost) == hipSuccess);
      hipFree(prio);
      hipFree(cdf);
      const std::uint8_t want[12] = {49, 24, 49, 24, 4, 24, 49, 24, 49, 59, 49, 59};
      EXPECT(level0 == 1);
      EXPECT(std::memcmp(hp, want, 12) == 0);
      EXPECT(hc[0] == 0 && hc[5] == 1 && hc[25] == 5 && hc[50] == 10 && hc[60] == 12 && hc[63] == 12);
      bool thrown = false;
      try
      {
        dt::compute_block_priorities(mask, nullptr, nullptr);
      }
      catch (const std::invalid_argument&)
      {
        thrown = true;
      }
      EXPECT(thrown);
    });
    run_test("Errors.InvalidLambda", [] {
      dt::image2d<std::uint8_t> img(8, 8, dt::e_memory_kind::GPU), mask(8, 8, dt::e_memory_kind::GPU);
      bool                      thrown = false;
      try
      {
        dt::generalised_distance_transform_task(img, mask, 1.5f);
      }
      catch (const std::invalid_argument&)
      {
        thrown = true;
      }
      EXPECT(thrown);
    });
  }
} // namespace

int main(int argc, char** argv)
{
  bool gpu = false;
  for (int i = 1; i < argc; ++i)
  {
    if (!std::strcmp(argv[i], "--gpu"))
      gpu = true;
    else if (!std::strcmp(argv[i], "--golden") && i + 1 < argc)
      g_golden = argv[++i];
  }
  cpu_tests();
  if (gpu)
    gpu_tests();
  std::printf("[==========] %d tests ran, %d failed\n", g_run, g_fail);
  return g_fail == 0 ? 0 : 1;
}
