// Input and output formats of the reconstruction: MRC projection stacks and
// volumes, tilt-angle files, TIFF images and image-stack directories
// (metadata.xml + one TIFF per projection).  Host code; the reconstruction
// itself never touches files.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "etn/etn.h"

namespace etn {

// MRC header as the reference reads and writes it (MRCHeader.h:3-113): 1024
// bytes, native (little) endian, `next` extended-header bytes follow it.
struct MrcHeader {
    int32_t nx, ny, nz, mode;
    int32_t nxstart, nystart, nzstart;
    int32_t mx, my, mz;
    float cell_dim[3];
    float cell_angle[3];
    int32_t mapc, mapr, maps;
    float dmin, dmax, dmean;
    int16_t ispg, nsymbt;
    int32_t next;
    int16_t creator_id;
    char extra1[30];
    int16_t nint, nreal;
    char extra2[20];
    int32_t imod_stamp, imod_flags;
    int16_t idtype, lens, nd1, nd2, vd1, vd2;
    float tilt_angles[6];
    float origin[3];
    char map[4];
    char machine_stamp[4];
    float rms;
    int32_t nlabel;
    char labels[10][80];
};
static_assert(sizeof(MrcHeader) == 1024, "MRC header must be 1024 bytes");

// IOEnumerations.h: voxel value types (equal to the MRC modes) and the
// coordinate orders a volume can be written in.
enum class VoxelValueType : int32_t { Unsigned8 = 0, Signed16 = 1, Float32 = 2, Unsigned16 = 6 };
enum class CoordinateOrder { XYZ, XZY, YXZ, YZX, ZXY, ZYX };

VoxelValueType parse_voxel_value_type(const std::string& s);     // IOEnumerations.cpp:40-61
CoordinateOrder parse_coordinate_order(const std::string& s);    // IOEnumerations.cpp:7-38
const char* voxel_value_type_name(VoxelValueType t);
const char* coordinate_order_name(CoordinateOrder o);

// A projection stack in memory: count images of nu x nv floats, stack order,
// with the tilt angles (degrees) and the untilted base geometry.
struct ProjectionStack {
    std::string kind;                 // "MRCStack" | "ImageStackDirectoryDatasource"
    uint32_t nu = 0, nv = 0, count = 0;
    std::vector<float> data;          // [count][nv][nu]
    std::vector<float> tilt_angles;   // count entries
    etn_scanner_geometry base{};
};

// ImageStackDatasource::readTiltAngles (ImageStackDatasource.cpp:40-73): one
// angle per whitespace-separated token; the count must equal `expected`.
std::vector<float> read_tilt_angles(const std::string& path, uint32_t expected);

// MRCStack (MRCStack.cpp:136-263): modes 0 (uint8), 1 (int16), 2 (float32),
// 6 (uint16); optional log(dataMax / (raw - min)) transform; base geometry
// source (-nx/2, -ny/2, nx), detector (-nx/2, -ny/2, -nx), pitches (1,0,0),
// (0,1,0).  Empty tilt_path gives zero angles.
ProjectionStack read_mrc_stack(const std::string& path, const std::string& tilt_path, bool logarithmize);

// ImageStackDirectoryDatasource (ImageStackDirectoryDatasource.cpp:181-276).
ProjectionStack read_image_stack_directory(const std::string& directory);

// ImageStackDatasourceFactory::instantiate (ImageStackDataSourceFactory.cpp):
// a directory holding metadata.xml first, then a regular .mrc / .ali file.
ProjectionStack open_projection_stack(const std::string& path, const std::string& tilt_path, bool logarithmize);

// ImageStackDirectoryDatasource::writeMetaDataFile plus one float TIFF per
// projection named projection_<i>.tif.
void write_image_stack_directory(const std::string& directory, const float* data, uint32_t count, uint32_t nu,
                                 uint32_t nv, const float* tilt_angles, const etn_scanner_geometry& base,
                                 const uint32_t volume_resolution[3]);

// TIFF: grey images, strips or tiles, compression none / LZW / PackBits /
// Deflate, predictors 1-3.  Samples convert as FreeImage's ConvertToFloat
// does: float32 unchanged, uint8 / 255, uint16 / 65535.  Row 0 of the file is
// image row y = 0 (ImageDeserializer.cpp:48-61 flips FreeImage's bottom-up rows).
uint32_t tiff_page_count(const std::string& path);
std::vector<float> read_tiff(const std::string& path, uint32_t page, uint32_t* width, uint32_t* height);
void write_tiff(const std::string& path, const float* data, uint32_t width, uint32_t height);

// VolumeFileFormatWriter::write + MRCWriter::writeHeader
// (VolumeFileFormatWriter.cpp:11-144, MRCWriter.cpp:43-80).  `volume` is
// [z][y][x] with res = (nx, ny, nz).
void write_mrc_volume(const std::string& path, const float* volume, const uint32_t res[3], VoxelValueType type,
                      CoordinateOrder order, bool invert);

// VolumeDeserializer::load (VolumeDeserializer.cpp) reads a volume file as an
// MRC stack without transform; BlockIterativeConfiguration::
// applyInitialVolumeToWorkingVolume (BlockIterativeConfiguration.cpp:156-196)
// maps the file's coordinates by `order`.  Returns [z][y][x] and its res.
std::vector<float> read_mrc_volume(const std::string& path, CoordinateOrder order, uint32_t res[3]);

// Host conversion of the reference's HalfFloatVolume (Float16Compressor.cpp:7-23):
// mantissa truncated toward zero, values above the largest half become infinity.
uint16_t float_to_half_truncate(float f);
float half_to_float(uint16_t h);
uint16_t float_to_bfloat16_rne(float f);

// Path helpers with boost::filesystem meaning: '\' becomes '/', relative paths
// are taken against `base_dir`, "." and ".." are folded.
std::string normalize_path(const std::string& p);
std::string join_path(const std::string& base_dir, const std::string& p);
std::string parent_path(const std::string& p);
bool is_directory(const std::string& p);
bool is_regular_file(const std::string& p);
std::string extension(const std::string& p);   // ".mrc" (with the dot) or ""

}  // namespace etn
