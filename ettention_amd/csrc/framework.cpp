// Command-line reconstruction (see framework.hpp).
#include "framework.hpp"

#include <sys/stat.h>

#include <cctype>
#include <chrono>
#include <cmath>
#include <cstring>
#include <cstdio>
#include <fstream>
#include <memory>
#include <sstream>

#include "block_iterative.hpp"
#include "wbp.hpp"
#include "common.hpp"

namespace etn {

namespace {

const char* kShortHelp =
    "Usage: etn --config <file.xml> [--<parameter> <value> ...]\n"
    "  -c, --config <file>          parameter file of the reconstruction\n"
    "  -e, --explicitDeviceId <id>  GPU to use (same as hardware.openclDeviceId)\n"
    "  -d, --devices                list the GPUs of this machine\n"
    "  -h, --help                   show this message\n"
    "  -H, --parameters             show every parameter\n";

const char* kParameters =
    "Parameters (XML <a><b value=\"...\"/></a> or --a.b <value>; the command line wins):\n"
    "  input.projections             MRC stack (.mrc/.ali) or directory with metadata.xml   [required]\n"
    "  input.tiltAngles              tilt file (.tlt), one angle in degrees per projection\n"
    "  input.xAxisTilt               fixed tilt of the geometry about x, degrees            [0]\n"
    "  input.logarithmize            MRC input: log(dataMax / (raw - min))                  [false]\n"
    "  input.skipProjections         'none' or 1-based list / ranges, e.g. '1,5-7'\n"
    "  output.filename               output volume                                          [required]\n"
    "  output.format                 mrc                                                    [from filename]\n"
    "  output.options.voxelType      unsigned8|signed16|unsigned16|float32                  [float32]\n"
    "  output.options.orientation    xzy_order|xyz_order|yxz_order|yzx_order|zxy_order|zyx_order [xzy_order]\n"
    "  output.options.invert         true|false                                             [false]\n"
    "  volume.dimension              size in world units 'x y z'                            [required]\n"
    "  volume.origin                 centre 'x y z'                                         [0 0 0]\n"
    "  volume.resolution             voxels 'x y z'                                         [dimension]\n"
    "  volume.initValue              constant start value                                   [0]\n"
    "  volume.initFile               start volume (MRC)\n"
    "  volume.initFileOrientation    coordinate order of volume.initFile                    [xzy_order]\n"
    "  algorithm                     sart|sirt|blockIterative                               [required]\n"
    "  algorithm.blockSize           projections per block (blockIterative)\n"
    "  algorithm.numberOfIterations  iterations                                             [required]\n"
    "  algorithm.lambda              relaxation                                             [required]\n"
    "  algorithm.useLongObjectCompensation  true|false                                      [false]\n"
    "  forwardProjection.samples     rays per pixel (a square)                              [1]\n"
    "  backProjection.samples        samples per voxel (a square)                           [1]\n"
    "  optimization.internalVoxelRepresentationType  float|half|bfloat16                    [float]\n"
    "  optimization.projectionIteration  random|identity|maxangle|buckets                   [random]\n"
    "  priorknowledge.maskVolumeFilename  uint8 mask volume (MRC, x fastest, z slowest)\n"
    "  hardware.subVolumeCount       z-slabs the volume is processed in                     [automatic]\n"
    "  hardware.openclDeviceId       GPU ordinal\n"
    "  debug.profilingLevel          none|overall|normal|detail|timing                      [overall]\n"
    "  debug.writeIntermediateVolumes                 volume after every block\n"
    "  debug.writeIntermediateVolumesAfterIterations  volume after every iteration\n"
    "  debug.infoPath                directory of the intermediate volumes                  [debug]\n";

double seconds_since(std::chrono::steady_clock::time_point t0)
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}

void make_directories(const std::string& p)
{
    std::string acc;
    std::stringstream ss(p);
    std::string part;
    if (!p.empty() && p[0] == '/') acc = "/";
    while (std::getline(ss, part, '/')) {
        if (part.empty()) continue;
        acc += (acc.empty() || acc.back() == '/' ? "" : "/") + part;
        ::mkdir(acc.c_str(), 0777);
    }
    if (!is_directory(p)) fail("cannot create directory " + p);
}

// OutputParameterSet::testAvailabilityOfOutputDirectory and
// VolumeSerializer::findWriter (VolumeSerializer.cpp:42-56).
void check_output(const ReconstructionConfig& c)
{
    {
        std::ofstream f(c.output_filename, std::ios::trunc);
        if (!f.good())
            fail("Cannot write to output path. Check its availability and/or permissions. Path: " + c.output_filename);
        f << "Test output folder availability.";
    }
    std::remove(c.output_filename.c_str());
    const std::string ext = c.output_format == ".mrc" ? ".mrc" : extension(c.output_filename);
    if (ext != ".mrc")
        unsupported("No suitable serializer for volume file format " +
                    (c.output_format.empty() ? ext : c.output_format) + ". Plugin missing?");
}

etn_voxel_type voxel_type_of(const std::string& r)
{
    if (r == "half") return ETN_VOXEL_FLOAT16;
    if (r == "bfloat16") return ETN_VOXEL_BFLOAT16;
    return ETN_VOXEL_FLOAT32;
}

// The working volume as float on the host, [z][y][x].
std::vector<float> download_volume(ReconstructionAlgorithm& op)
{
    GpuVolume* v = op.getReconstructedVolume();
    op.synchronize();
    const size_t n = v->properties().voxel_count();
    std::vector<float> out(n);
    if (v->voxel_type() == ETN_VOXEL_FLOAT32) {
        ETN_HIP(hipMemcpy(out.data(), v->data(), n * 4, hipMemcpyDeviceToHost));
        return out;
    }
    std::vector<uint16_t> raw(n);
    ETN_HIP(hipMemcpy(raw.data(), v->data(), n * 2, hipMemcpyDeviceToHost));
    if (v->voxel_type() == ETN_VOXEL_FLOAT16) {
        for (size_t i = 0; i < n; ++i) out[i] = half_to_float(raw[i]);
    } else {
        for (size_t i = 0; i < n; ++i) {
            const uint32_t u = (uint32_t)raw[i] << 16;
            std::memcpy(&out[i], &u, 4);
        }
    }
    return out;
}

int profiling_rank(const std::string& level)
{
    if (level == "none") return 0;
    if (level == "overall") return 1;
    if (level == "normal") return 2;
    if (level == "detail") return 3;
    return 4;
}

}  // namespace

void run_reconstruction(const ReconstructionConfig& c, std::ostream& log)
{
    check_output(c);
    const bool dumps = c.write_intermediate_volumes || c.write_intermediate_volumes_after_iterations;
    if (dumps) make_directories(c.debug_info_path);

    if (c.device_id >= 0) ETN_HIP(hipSetDevice(c.device_id));

    ProjectionStack stack = open_projection_stack(c.projections, c.tilt_angles, c.logarithmize);
    log << "Chosen ImageStackDatasource: " << stack.kind << ".\n";
    log << "Projections: " << stack.count << " of " << stack.nu << " x " << stack.nv << " pixels\n";

    const uint32_t* res = c.resolution;
    const size_t voxels = (size_t)res[0] * res[1] * res[2];
    if (voxels == 0) fail("volume resolution must be positive in every axis");

    DeviceBuffer projections((size_t)stack.count * stack.nu * stack.nv * sizeof(float));
    ETN_HIP(hipMemcpy(projections.get(), stack.data.data(), projections.bytes(), hipMemcpyHostToDevice));

    // BlockIterativeConfiguration.cpp:45-56: the mask file is read as an MRC
    // stack straight into a byte volume (x fastest, z slowest), values truncated.
    DeviceBuffer mask;
    if (!c.mask_file.empty()) {
        uint32_t mres[3];
        std::vector<float> m = read_mrc_volume(c.mask_file, CoordinateOrder::XYZ, mres);
        if (mres[0] != res[0] || mres[1] != res[1] || mres[2] != res[2])
            fail("Resolution of mask volume " + c.mask_file + " does not match resolution given in config file!");
        std::vector<uint8_t> bytes(voxels);
        for (size_t i = 0; i < voxels; ++i) {
            const float v = m[i];
            bytes[i] = std::isnan(v) || v <= 0.0f ? 0 : v >= 255.0f ? 255 : (uint8_t)v;
        }
        mask.allocate(voxels);
        ETN_HIP(hipMemcpy(mask.get(), bytes.data(), voxels, hipMemcpyHostToDevice));
    }

    std::string id = c.algorithm;
    for (char& ch : id) ch = (char)std::tolower((unsigned char)ch);
    const bool wbp = id == "wbp";
    etn_reconstruction_desc d{};
    d.algorithm = "blockiterative";   // the block size is already resolved (AlgorithmParameterSet)
    d.block_size = c.block_size;
    d.iterations = c.iterations;
    d.lambda = c.lambda;
    d.use_long_object_compensation = c.use_long_object_compensation ? 1 : 0;
    d.fp_samples = c.fp_samples;
    d.bp_samples = c.bp_samples;
    d.projection_iteration = c.projection_iteration.c_str();
    d.skip_projections = c.skip_projections.empty() ? nullptr : c.skip_projections.data();
    d.skip_count = (uint32_t)c.skip_projections.size();
    // WBP always works on a FloatVolume (WBP.cpp:20)
    d.voxel_type = wbp ? ETN_VOXEL_FLOAT32 : voxel_type_of(c.voxel_representation);
    for (int k = 0; k < 3; ++k) {
        const float half = 0.5f * c.dimension[k];   // VolumeParameterSet::getBoundingBox
        d.volume.resolution[k] = res[k];
        d.volume.bbox_min[k] = c.origin[k] - half;
        d.volume.bbox_max[k] = c.origin[k] + half;
    }
    d.volume.subvolume_count = c.subvolume_count;
    d.init_value = c.init_value;
    d.base_geometry = stack.base;
    d.projection_count = stack.count;
    d.tilt_angles = stack.tilt_angles.data();
    d.x_axis_tilt = c.x_axis_tilt;
    d.projections = projections.as<float>();
    d.mask = mask.as<uint8_t>();
    d.flags = c.write_intermediate_volumes ? (ETN_RECON_FLAG_GENERIC_KERNELS | ETN_RECON_FLAG_NO_GRAPH) : 0u;

    std::unique_ptr<ReconstructionAlgorithm> algorithm;
    BlockIterativeReconstructionOperator* block_iterative = nullptr;
    if (wbp) {
        algorithm = std::make_unique<WeightedBackProjection>(d, nullptr);
    } else {
        auto bi = std::make_unique<BlockIterativeReconstructionOperator>(d, nullptr);
        block_iterative = bi.get();
        algorithm = std::move(bi);
    }
    ReconstructionAlgorithm& op = *algorithm;

    if (!c.init_file.empty()) {
        // VolumeDeserializer::load + applyInitialVolumeToWorkingVolume; a half
        // working volume stores the values through the host compressor.
        uint32_t ires[3];
        std::vector<float> init = read_mrc_volume(c.init_file, c.init_file_orientation, ires);
        if (ires[0] != res[0] || ires[1] != res[1] || ires[2] != res[2])
            fail("Resolution of volume from file " + c.init_file + " does not match resolution given in config file!");
        GpuVolume* v = op.getReconstructedVolume();
        std::vector<uint16_t> packed;
        const void* src = init.data();
        if (d.voxel_type != ETN_VOXEL_FLOAT32) {
            packed.resize(voxels);
            for (size_t i = 0; i < voxels; ++i)
                packed[i] = d.voxel_type == ETN_VOXEL_FLOAT16 ? float_to_half_truncate(init[i])
                                                              : float_to_bfloat16_rne(init[i]);
            src = packed.data();
        }
        ETN_HIP(hipStreamSynchronize(op.stream()));
        ETN_HIP(hipMemcpy(v->data(), src, v->bytes(), hipMemcpyHostToDevice));
        log << "Initial volume read from " << c.init_file << "\n";
    }

    const int prof = profiling_rank(c.profiling_level);
    uint32_t iteration = 0;
    // intermediate volumes are float32 in x-y-z order (default OutputParameterSet,
    // BlockIterativeReconstructionAlgorithmLogger.cpp:170-176)
    auto dump = [&](const std::string& name) {
        std::vector<float> v = download_volume(op);
        write_mrc_volume(join_path(c.debug_info_path, name), v.data(), res, VoxelValueType::Float32,
                         CoordinateOrder::XYZ, false);
    };
    if (c.write_intermediate_volumes && block_iterative) {
        block_iterative->setBlockEndCallback([&](uint32_t block) {
            char name[64];
            std::snprintf(name, sizeof name, "intermediate_i%03u_b%03u.mrc", iteration, block);
            dump(name);
        });
    }

    log << "Launching reconstruction algorithm...\n";
    const auto t0 = std::chrono::steady_clock::now();
    if (wbp) {
        // WBP::run filters and back-projects every projection once; the
        // iteration count and the intermediate volumes do not apply
        op.run();
    }
    for (iteration = 0; !wbp && iteration < c.iterations; ++iteration) {
        const auto ti = std::chrono::steady_clock::now();
        op.runNiterations(1);
        if (prof >= 2) {
            op.synchronize();
            log << "Time spent for iteration " << iteration << ": " << seconds_since(ti) << "s.\n";
        }
        if (c.write_intermediate_volumes_after_iterations) {
            char name[64];
            std::snprintf(name, sizeof name, "intermediate_i%03u.mrc", iteration);
            dump(name);
        }
    }
    op.synchronize();
    if (prof >= 1) log << "Total time spent for reconstruction: " << seconds_since(t0) << "s.\n";

    std::vector<float> volume = download_volume(op);
    write_mrc_volume(c.output_filename, volume.data(), res, c.output_voxel_type, c.output_orientation,
                     c.output_invert);
    log << "Reconstruction finished.\n";
}

ReconstructionConfig resolve_config(int argc, const char* const* argv)
{
    auto cli = std::make_shared<CommandLineParameterSource>(argc, argv);
    CascadingParameterSource all;
    if (cli->exists("config")) all.add_source(std::make_shared<XmlParameterSource>(cli->string("config")));
    all.add_source(cli);   // main.cpp:36-49: the command line is added last and wins
    return ReconstructionConfig::from(all);
}

int framework_main(int argc, const char* const* argv, std::ostream& log)
{
    try {
        log << "Ettention (MI355X / gfx950 reconstruction library)\n";
        CommandLineParameterSource cli(argc, argv);
        switch (cli.mode()) {
            case CommandLineParameterSource::Mode::Help:
                log << kShortHelp;
                return 0;
            case CommandLineParameterSource::Mode::HelpAndParameters:
                log << kShortHelp << kParameters;
                return 0;
            case CommandLineParameterSource::Mode::Devices: {
                int n = 0;
                ETN_HIP(hipGetDeviceCount(&n));
                for (int i = 0; i < n; ++i) {
                    hipDeviceProp_t p;
                    ETN_HIP(hipGetDeviceProperties(&p, i));
                    log << i << ": " << p.name << " (" << p.gcnArchName << "), " << p.multiProcessorCount
                        << " compute units, " << (p.totalGlobalMem >> 30) << " GiB\n";
                }
                return 0;
            }
            default:
                break;
        }
        const bool with_config = cli.mode() == CommandLineParameterSource::Mode::Reconstruct;
        if (with_config) log << "Loading configuration from XML file...\n";
        try {
            const ReconstructionConfig c = resolve_config(argc, argv);
            run_reconstruction(c, log);
        } catch (const ParameterNotFound& e) {
            if (with_config) throw;
            log << "Error: Missing required parameter " << e.key << ". Set it by using --" << e.key
                << " <value> cmd argument or load from XML config file by --config <path> argument\n";
        }
        return 0;
    } catch (const std::exception& e) {
        log << "Error: " << e.what() << "\n";
        return 1;
    } catch (...) {
        log << "Error: Unknown error\n";
        return 2;
    }
}

}  // namespace etn
