// extern "C" entry points of the file formats and the command-line
// reconstruction (include/etn/etn.h).  Exceptions become etn_status +
// etn_last_error(), as in capi.cpp.
#include <cstring>
#include <iostream>
#include <memory>
#include <string>

#include "common.hpp"
#include "config.hpp"
#include "etn/etn.h"
#include "framework.hpp"
#include "io.hpp"

namespace {

template <typename F>
etn_status guarded_io(F&& f)
{
    try {
        f();
        return ETN_OK;
    } catch (const etn::Error& e) {
        return etn::record_error(e.status, e.what());
    } catch (const etn::ParameterNotFound& e) {
        return etn::record_error(ETN_ERROR_INVALID_ARGUMENT, e.what());
    } catch (const std::bad_alloc&) {
        return etn::record_error(ETN_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (const std::exception& e) {
        return etn::record_error(ETN_ERROR_INTERNAL, e.what());
    } catch (...) {
        return etn::record_error(ETN_ERROR_INTERNAL, "unknown error");
    }
}

void need_io(const void* p, const char* what)
{
    if (!p) etn::fail(std::string(what) + " is null");
}

}  // namespace

struct etn_projection_stack {
    etn::ProjectionStack stack;
};

extern "C" {

etn_status etn_projection_stack_open(const char* projections, const char* tilt_angles, int logarithmize,
                                     etn_projection_stack** out)
{
    return guarded_io([&] {
        need_io(projections, "projections path");
        need_io(out, "output handle");
        *out = nullptr;
        auto s = std::make_unique<etn_projection_stack>();
        s->stack = etn::open_projection_stack(etn::normalize_path(projections),
                                              tilt_angles && *tilt_angles ? etn::normalize_path(tilt_angles) : "",
                                              logarithmize != 0);
        *out = s.release();
    });
}

etn_status etn_projection_stack_destroy(etn_projection_stack* s)
{
    return guarded_io([&] { delete s; });
}

etn_status etn_projection_stack_info(const etn_projection_stack* s, uint32_t* nu, uint32_t* nv, uint32_t* count,
                                     etn_scanner_geometry* base)
{
    return guarded_io([&] {
        need_io(s, "projection stack");
        if (nu) *nu = s->stack.nu;
        if (nv) *nv = s->stack.nv;
        if (count) *count = s->stack.count;
        if (base) *base = s->stack.base;
    });
}

etn_status etn_projection_stack_read(const etn_projection_stack* s, float* projections, float* tilt_angles)
{
    return guarded_io([&] {
        need_io(s, "projection stack");
        if (projections) std::memcpy(projections, s->stack.data.data(), s->stack.data.size() * sizeof(float));
        if (tilt_angles)
            std::memcpy(tilt_angles, s->stack.tilt_angles.data(), s->stack.tilt_angles.size() * sizeof(float));
    });
}

etn_status etn_image_stack_directory_write(const char* directory, const float* projections, uint32_t count,
                                           uint32_t nu, uint32_t nv, const float* tilt_angles,
                                           const etn_scanner_geometry* base, const uint32_t volume_resolution[3])
{
    return guarded_io([&] {
        need_io(directory, "directory");
        need_io(projections, "projections");
        need_io(tilt_angles, "tilt angles");
        need_io(base, "base geometry");
        need_io(volume_resolution, "volume resolution");
        if (!count || !nu || !nv) etn::fail("empty projection stack");
        etn::write_image_stack_directory(directory, projections, count, nu, nv, tilt_angles, *base,
                                         volume_resolution);
    });
}

etn_status etn_tiff_read(const char* path, uint32_t page, uint32_t* width, uint32_t* height, float* pixels)
{
    return guarded_io([&] {
        need_io(path, "path");
        uint32_t w = 0, h = 0;
        std::vector<float> img = etn::read_tiff(path, page, &w, &h);
        if (width) *width = w;
        if (height) *height = h;
        if (pixels) std::memcpy(pixels, img.data(), img.size() * sizeof(float));
    });
}

etn_status etn_tiff_page_count(const char* path, uint32_t* pages)
{
    return guarded_io([&] {
        need_io(path, "path");
        need_io(pages, "output");
        *pages = etn::tiff_page_count(path);
    });
}

etn_status etn_tiff_write(const char* path, const float* pixels, uint32_t width, uint32_t height)
{
    return guarded_io([&] {
        need_io(path, "path");
        need_io(pixels, "pixels");
        etn::write_tiff(path, pixels, width, height);
    });
}

etn_status etn_mrc_write_volume(const char* path, const float* volume, const uint32_t resolution[3],
                                const char* voxel_type, const char* orientation, int invert)
{
    return guarded_io([&] {
        need_io(path, "path");
        need_io(volume, "volume");
        need_io(resolution, "resolution");
        etn::write_mrc_volume(path, volume, resolution,
                              etn::parse_voxel_value_type(voxel_type ? voxel_type : "float32"),
                              etn::parse_coordinate_order(orientation ? orientation : "xzy_order"), invert != 0);
    });
}

etn_status etn_mrc_read_volume(const char* path, const char* orientation, uint32_t resolution[3], float* volume)
{
    return guarded_io([&] {
        need_io(path, "path");
        need_io(resolution, "resolution");
        std::vector<float> v =
            etn::read_mrc_volume(path, etn::parse_coordinate_order(orientation ? orientation : "xzy_order"),
                                 resolution);
        if (volume) std::memcpy(volume, v.data(), v.size() * sizeof(float));
    });
}

etn_status etn_config_resolve(int argc, const char* const* argv, char* json, size_t capacity, size_t* length)
{
    return guarded_io([&] {
        need_io(argv, "argv");
        const std::string s = etn::resolve_config(argc, argv).to_json();
        if (length) *length = s.size();
        if (json && capacity) {
            const size_t n = std::min(capacity - 1, s.size());
            std::memcpy(json, s.data(), n);
            json[n] = '\0';
        }
    });
}

int etn_main(int argc, const char* const* argv)
{
    const int rc = etn::framework_main(argc, argv, std::cout);
    std::cout.flush();
    return rc;
}

}  // extern "C"
