// C++ API of include/dt/ on top of the C ABI of include/dtk.h.
//
// Every call is synchronous like the reference's (which ends with
// cudaDeviceSynchronize, dt/src/generalised_distance_transform.cu:133).
// Status codes become exceptions: DTK_ERR_INVALID_ARGUMENT ->
// std::invalid_argument, everything else -> std::runtime_error
// (dt/src/generalised_distance_transform.cu:136-137).
#include <hip/hip_runtime.h>

#include <stdexcept>
#include <string>
#include <type_traits>

#include <dt/block_priority.hpp>
#include <dt/generalised_distance_transform.hpp>
#include <dt/transfert.hpp>
#include <dtk.h>

namespace dt
{
  namespace
  {
    void check(dtk_status s, const char* what)
    {
      if (s == DTK_OK)
        return;
      const std::string msg = std::string(what) + ": " + dtk_last_error();
      if (s == DTK_ERR_INVALID_ARGUMENT)
        throw std::invalid_argument(msg);
      throw std::runtime_error(msg);
    }

    template <typename T>
    constexpr int dtype_of()
    {
      if constexpr (std::is_same_v<T, std::uint8_t>)
        return DTK_U8;
      else if constexpr (std::is_same_v<T, rgb8>)
        return DTK_RGB8;
      else
      {
        static_assert(std::is_same_v<T, float>, "unsupported pixel type");
        return DTK_F32;
      }
    }

    template <typename T>
    dtk_image to_dtk(const image2d_view<T>& v)
    {
      dtk_image d;
      d.data   = const_cast<std::uint8_t*>(v.buffer());
      d.width  = v.width();
      d.height = v.height();
      d.pitch  = v.pitch();
      d.dtype  = dtype_of<T>();
      return d;
    }

    void require_gpu(const image2d_view<void>& v, const char* name)
    {
      if (v.memory_kind() != e_memory_kind::GPU)
        throw std::invalid_argument(std::string(name) + " must be a GPU image");
    }

    dtk_params to_params(float lambda, float v, const ggdt_options& o)
    {
      dtk_params p;
      dtk_params_default(&p);
      p.lambda           = lambda;
      p.v                = v;
      p.connectivity     = o.connectivity == 4 ? DTK_C4 : DTK_C8;
      p.schedule         = o.schedule == e_schedule::raster   ? DTK_SCHED_RASTER
                           : o.schedule == e_schedule::blocks ? DTK_SCHED_BLOCKS
                                                              : DTK_SCHED_TASK;
      p.max_iterations   = o.max_iterations;
      p.tile_rows        = o.tile_rows;
      p.max_sweeps       = o.max_sweeps;
      p.waves_per_cu     = o.waves_per_cu;
      p.round_per_blocks = o.round_per_blocks;
      return p;
    }

    void run(const dtk_image& img, const dtk_image& mask, dtk_image& dist, const dtk_params& p)
    {
      dtk_stats st{};
      check(dtk_ggdt(&img, &mask, &dist, &p, nullptr, &st), "generalised_distance_transform");
      if (hipStreamSynchronize(nullptr) != hipSuccess)
        throw std::runtime_error("Error while running distance transform");
    }

    template <typename TI, typename TM>
    void checked(const image2d_view<TI>& img, const image2d_view<TM>& mask, const image2d_view<float>& dist)
    {
      if (img.width() != dist.width() || img.height() != dist.height() || img.width() != mask.width() ||
          img.height() != mask.height())
        throw std::invalid_argument("image, mask and distance sizes differ");
      require_gpu(img, "image");
      require_gpu(mask, "mask");
      require_gpu(dist, "distance");
    }

    void ggdt_u8(const image2d_view<std::uint8_t>& img, const image2d_view<std::uint8_t>& mask,
                 image2d_view<float>& dist, float lambda, float v, e_schedule sched, int* rpb)
    {
      checked(img, mask, dist);
      ggdt_options o;
      o.schedule         = sched;
      o.round_per_blocks = rpb;
      dtk_image  di = to_dtk(img), dm = to_dtk(mask), dd = to_dtk(dist);
      dtk_params p = to_params(lambda, v, o);
      run(di, dm, dd, p);
    }
  } // namespace

  void host_to_device(const image2d_view<void>& src, image2d_view<void>& dst)
  {
    if (src.memory_kind() != e_memory_kind::CPU || dst.memory_kind() != e_memory_kind::GPU)
      throw std::invalid_argument("host_to_device needs a CPU source and a GPU destination");
    if (src.width() != dst.width() || src.height() != dst.height() || src.elem_size() != dst.elem_size())
      throw std::invalid_argument("host_to_device between different shapes");
    check(dtk_copy2d(dst.buffer(), dst.pitch(), src.buffer(), src.pitch(), (long long)src.width() * src.elem_size(),
                     src.height(), 1),
          "host_to_device");
  }

  void device_to_host(const image2d_view<void>& src, image2d_view<void>& dst)
  {
    if (src.memory_kind() != e_memory_kind::GPU || dst.memory_kind() != e_memory_kind::CPU)
      throw std::invalid_argument("device_to_host needs a GPU source and a CPU destination");
    if (src.width() != dst.width() || src.height() != dst.height() || src.elem_size() != dst.elem_size())
      throw std::invalid_argument("device_to_host between different shapes");
    check(dtk_copy2d(dst.buffer(), dst.pitch(), src.buffer(), src.pitch(), (long long)src.width() * src.elem_size(),
                     src.height(), 2),
          "device_to_host");
  }

  void generalised_distance_transform(const image2d_view<std::uint8_t>& img, const image2d_view<std::uint8_t>& mask,
                                      image2d_view<float>& dist, float lambda, float v)
  {
    ggdt_u8(img, mask, dist, lambda, v, e_schedule::raster, nullptr);
  }

  image2d<float> generalised_distance_transform(const image2d_view<std::uint8_t>& img,
                                                const image2d_view<std::uint8_t>& mask, float lambda, float v)
  {
    image2d<float> D(img.width(), img.height(), e_memory_kind::GPU);
    generalised_distance_transform(img, mask, D, lambda, v);
    return D;
  }

  void generalised_distance_transform_blocks(const image2d_view<std::uint8_t>& img,
                                             const image2d_view<std::uint8_t>& mask, image2d_view<float>& dist,
                                             float lambda, float v, int* round_per_blocks)
  {
    ggdt_u8(img, mask, dist, lambda, v, e_schedule::blocks, round_per_blocks);
  }

  image2d<float> generalised_distance_transform_blocks(const image2d_view<std::uint8_t>& img,
                                                       const image2d_view<std::uint8_t>& mask, float lambda, float v,
                                                       int* round_per_blocks)
  {
    image2d<float> D(img.width(), img.height(), e_memory_kind::GPU);
    generalised_distance_transform_blocks(img, mask, D, lambda, v, round_per_blocks);
    return D;
  }

  void generalised_distance_transform_task(const image2d_view<std::uint8_t>& img,
                                           const image2d_view<std::uint8_t>& mask, image2d_view<float>& D, float lambda,
                                           float v)
  {
    ggdt_u8(img, mask, D, lambda, v, e_schedule::task, nullptr);
  }

  image2d<float> generalised_distance_transform_task(const image2d_view<std::uint8_t>& img,
                                                     const image2d_view<std::uint8_t>& mask, float lambda, float v)
  {
    image2d<float> D(img.width(), img.height(), e_memory_kind::GPU);
    generalised_distance_transform_task(img, mask, D, lambda, v);
    return D;
  }

  ggdt_options with_connectivity(ggdt_options o, const c4_t&)
  {
    o.connectivity = 4;
    return o;
  }
  ggdt_options with_connectivity(ggdt_options o, const c8_t&)
  {
    o.connectivity = 8;
    return o;
  }

  template <typename TI, typename TM>
  void generalised_distance_transform(const image2d_view<TI>& img, const image2d_view<TM>& mask,
                                      image2d_view<float>& dist, float lambda, float v, const ggdt_options& opt)
  {
    checked(img, mask, dist);
    dtk_image  di = to_dtk(img), dm = to_dtk(mask), dd = to_dtk(dist);
    dtk_params p = to_params(lambda, v, opt);
    run(di, dm, dd, p);
  }

  template <typename TI, typename TM>
  void generalised_distance_transform_batched(const std::vector<image2d_view<TI>>& imgs,
                                              const std::vector<image2d_view<TM>>& masks,
                                              std::vector<image2d_view<float>>& dists, float lambda, float v,
                                              const ggdt_options& opt)
  {
    if (imgs.size() != masks.size() || imgs.size() != dists.size())
      throw std::invalid_argument("batch vectors differ in length");
    std::vector<dtk_image> di, dm, dd;
    for (std::size_t i = 0; i < imgs.size(); ++i)
    {
      checked(imgs[i], masks[i], dists[i]);
      di.push_back(to_dtk(imgs[i]));
      dm.push_back(to_dtk(masks[i]));
      dd.push_back(to_dtk(dists[i]));
    }
    ggdt_options o = opt;
    o.schedule     = e_schedule::task;
    dtk_params p   = to_params(lambda, v, o);
    dtk_stats  st{};
    check(dtk_ggdt_batched((int)imgs.size(), di.data(), dm.data(), dd.data(), &p, nullptr, &st),
          "generalised_distance_transform_batched");
    if (hipStreamSynchronize(nullptr) != hipSuccess)
      throw std::runtime_error("Error while running distance transform");
  }

  template <typename TI, typename TM>
  void generalised_distance_transform_predecessors(const image2d_view<TI>& img, const image2d_view<TM>& mask,
                                                   const image2d_view<float>& dist, image2d_view<std::uint8_t>& pred,
                                                   float lambda, float v, const ggdt_options& opt)
  {
    checked(img, mask, dist);
    require_gpu(pred, "pred");
    dtk_image  di = to_dtk(img), dm = to_dtk(mask), dd = to_dtk(dist), dp = to_dtk(pred);
    dtk_params p = to_params(lambda, v, opt);
    check(dtk_ggdt_pred(&di, &dm, &dd, &dp, &p, nullptr), "generalised_distance_transform_predecessors");
    if (hipStreamSynchronize(nullptr) != hipSuccess)
      throw std::runtime_error("Error while computing predecessors");
  }

#define DT_INSTANTIATE(TI, TM)                                                                                        \
  template void generalised_distance_transform<TI, TM>(const image2d_view<TI>&, const image2d_view<TM>&,             \
                                                       image2d_view<float>&, float, float, const ggdt_options&);     \
  template void generalised_distance_transform_batched<TI, TM>(                                                      \
      const std::vector<image2d_view<TI>>&, const std::vector<image2d_view<TM>>&, std::vector<image2d_view<float>>&, \
      float, float, const ggdt_options&);                                                                            \
  template void generalised_distance_transform_predecessors<TI, TM>(                                                 \
      const image2d_view<TI>&, const image2d_view<TM>&, const image2d_view<float>&, image2d_view<std::uint8_t>&,    \
      float, float, const ggdt_options&);

  DT_INSTANTIATE(std::uint8_t, std::uint8_t)
  DT_INSTANTIATE(std::uint8_t, float)
  DT_INSTANTIATE(rgb8, std::uint8_t)
  DT_INSTANTIATE(rgb8, float)
  DT_INSTANTIATE(float, std::uint8_t)
  DT_INSTANTIATE(float, float)
  std::uint32_t compute_block_priorities(const image2d_view<std::uint8_t>& mask, std::uint8_t* priorities,
                                         std::uint32_t* cdf)
  {
    if (mask.memory_kind() != e_memory_kind::GPU)
      throw std::invalid_argument("compute_block_priorities: the mask must be a GPU image");
    dtk_image     dm     = to_dtk(mask);
    std::uint32_t level0 = 0;
    check(dtk_block_priorities(&dm, priorities, cdf, &level0, nullptr), "compute_block_priorities");
    return level0;
  }
} // namespace dt
