// fz/fzm_format.h — FZM container: on-disk structs, enums and helpers.
//
// Byte layout identical to the reference's include/fzm_format.h (so files
// interoperate): FZMHeaderCore (80 B) at offset 0, then FZMStageInfo x
// num_stages (256 B each), FZMBufferEntry x num_buffers (256 B each), then
// the payload (the pipeline outputs back to back, each data_size bytes).
// StageType / DataType numbering follows include/fzm_format.h:76-114.
#pragma once

#include <cstddef>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>

namespace fz {

constexpr uint32_t FZM_MAGIC = 0x464D5A32;  // first 4 file bytes, little-endian
constexpr uint8_t FZM_VERSION_MAJOR = 3;
constexpr uint8_t FZM_VERSION_MINOR = 1;
constexpr uint16_t FZM_VERSION = (uint16_t)((FZM_VERSION_MAJOR << 8) | FZM_VERSION_MINOR);
constexpr size_t FZM_LEGACY_HEADER_CORE_SIZE = 72;
constexpr uint16_t FZM_FLAG_HAS_DATA_CHECKSUM = 0x0001u;
constexpr uint16_t FZM_FLAG_HAS_HEADER_CHECKSUM = 0x0002u;
constexpr size_t FZM_MAX_BUFFERS = 32;
constexpr size_t FZM_MAX_NAME_LEN = 64;
constexpr size_t FZM_STAGE_CONFIG_SIZE = 128;
constexpr size_t FZM_MAX_SOURCES = 4;
constexpr size_t FZM_MAX_STAGE_INPUTS = 8;
constexpr size_t FZM_MAX_STAGE_OUTPUTS = 8;

// Raw on-disk version values up to 0xFF are bare major numbers.
constexpr uint8_t fzmVersionMajor(uint16_t v) { return v <= 0xFF ? (uint8_t)v : (uint8_t)(v >> 8); }
constexpr uint8_t fzmVersionMinor(uint16_t v) { return v <= 0xFF ? (uint8_t)0 : (uint8_t)(v & 0xFF); }

enum class StageType : uint16_t {
    UNKNOWN = 0,
    LORENZO_QUANT = 1,
    DIFFERENCE = 2,
    SCALE = 3,
    PASSTHROUGH = 4,
    RLE = 5,
    HUFFMAN = 6,
    BITPACK = 7,
    SPLIT = 10,
    MERGE = 11,
    LORENZO = 12,
    QUANTIZER = 14,
    ZIGZAG = 15,
    NEGABINARY = 16,
    BITSHUFFLE = 17,
    RZE = 18,
};

enum class DataType : uint8_t {
    UINT8 = 0,
    UINT16 = 1,
    UINT32 = 2,
    UINT64 = 3,
    INT8 = 4,
    INT16 = 5,
    INT32 = 6,
    INT64 = 7,
    FLOAT32 = 8,
    FLOAT64 = 9,
    UNKNOWN = 0xFF,
};

// Per-stage record (256 B).  Unused buffer ids are 0xFFFF.
struct FZMStageInfo {
    StageType stage_type;
    uint16_t stage_version;
    uint8_t num_inputs;
    uint8_t num_outputs;
    uint16_t reserved1;
    uint16_t input_buffer_ids[FZM_MAX_STAGE_INPUTS];
    uint16_t output_buffer_ids[FZM_MAX_STAGE_OUTPUTS];
    uint8_t stage_config[FZM_STAGE_CONFIG_SIZE];
    uint32_t config_size;
    uint8_t reserved2[84];

    FZMStageInfo() {
        std::memset(this, 0, sizeof(*this));
        stage_type = StageType::UNKNOWN;
        std::memset(input_buffer_ids, 0xFF, sizeof(input_buffer_ids));
        std::memset(output_buffer_ids, 0xFF, sizeof(output_buffer_ids));
    }
};
static_assert(sizeof(FZMStageInfo) == 256, "FZMStageInfo must be 256 bytes");

// Per-buffer record (256 B).  The 6 tail bytes after reserved2 are padding;
// this build always writes them as zero.
struct FZMBufferEntry {
    StageType stage_type;
    uint16_t stage_version;
    DataType data_type;
    uint8_t producer_output_idx;
    uint16_t dag_buffer_id;
    char name[FZM_MAX_NAME_LEN];
    uint64_t data_size;
    uint64_t allocated_size;
    uint64_t uncompressed_size;
    uint64_t byte_offset;
    uint8_t stage_config[FZM_STAGE_CONFIG_SIZE];
    uint32_t config_size;
    uint8_t reserved2[14];

    FZMBufferEntry() {
        std::memset(this, 0, sizeof(*this));
        stage_type = StageType::UNKNOWN;
        data_type = DataType::UINT8;
        dag_buffer_id = 0xFFFF;
    }
};
static_assert(sizeof(FZMBufferEntry) == 256, "FZMBufferEntry must be 256 bytes");
static_assert(offsetof(FZMBufferEntry, data_size) == 72, "FZMBufferEntry layout");
static_assert(offsetof(FZMBufferEntry, stage_config) == 104, "FZMBufferEntry layout");
static_assert(offsetof(FZMBufferEntry, config_size) == 232, "FZMBufferEntry layout");

// File header core (80 B).
struct FZMHeaderCore {
    uint32_t magic;
    uint16_t version;
    uint16_t num_buffers;
    uint64_t uncompressed_size;
    uint64_t compressed_size;
    uint64_t header_size;
    uint32_t num_stages;
    uint16_t num_sources;
    uint16_t flags;
    uint64_t source_uncompressed_sizes[FZM_MAX_SOURCES];
    uint32_t data_checksum;
    uint32_t header_checksum;

    FZMHeaderCore() {
        std::memset(this, 0, sizeof(*this));
        magic = FZM_MAGIC;
        version = FZM_VERSION;
        header_size = sizeof(FZMHeaderCore);
    }
    uint64_t computeHeaderSize() const {
        return sizeof(FZMHeaderCore) + (uint64_t)num_stages * sizeof(FZMStageInfo) +
               (uint64_t)num_buffers * sizeof(FZMBufferEntry);
    }
};
static_assert(sizeof(FZMHeaderCore) == 80, "FZMHeaderCore must be 80 bytes");
static_assert(offsetof(FZMHeaderCore, source_uncompressed_sizes) == 40, "FZMHeaderCore layout");
static_assert(offsetof(FZMHeaderCore, data_checksum) == 72, "FZMHeaderCore layout");

inline size_t getDataTypeSize(DataType t) {
    switch (t) {
    case DataType::UINT8: case DataType::INT8: return 1;
    case DataType::UINT16: case DataType::INT16: return 2;
    case DataType::UINT32: case DataType::INT32: case DataType::FLOAT32: return 4;
    case DataType::UINT64: case DataType::INT64: case DataType::FLOAT64: return 8;
    default: throw std::runtime_error("Unknown data type");
    }
}

inline std::string dataTypeToString(DataType t) {
    switch (t) {
    case DataType::UINT8: return "uint8";
    case DataType::UINT16: return "uint16";
    case DataType::UINT32: return "uint32";
    case DataType::UINT64: return "uint64";
    case DataType::INT8: return "int8";
    case DataType::INT16: return "int16";
    case DataType::INT32: return "int32";
    case DataType::INT64: return "int64";
    case DataType::FLOAT32: return "float32";
    case DataType::FLOAT64: return "float64";
    default: return "unknown";
    }
}

inline std::string stageTypeToString(StageType t) {
    switch (t) {
    case StageType::LORENZO_QUANT: return "LorenzoQuant";
    case StageType::DIFFERENCE: return "Difference";
    case StageType::SCALE: return "Scale";
    case StageType::PASSTHROUGH: return "PassThrough";
    case StageType::RLE: return "RLE";
    case StageType::HUFFMAN: return "Huffman";
    case StageType::BITPACK: return "BitPack";
    case StageType::SPLIT: return "Split";
    case StageType::MERGE: return "Merge";
    case StageType::QUANTIZER: return "Quantizer";
    case StageType::ZIGZAG: return "Zigzag";
    case StageType::NEGABINARY: return "Negabinary";
    case StageType::BITSHUFFLE: return "Bitshuffle";
    case StageType::RZE: return "RZE";
    case StageType::LORENZO: return "Lorenzo";
    default: return "Unknown";
    }
}

// IEEE 802.3 CRC32 (reflected, poly 0xEDB88320); `crc` is the running
// register: start with 0xFFFFFFFF, finish with ^ 0xFFFFFFFF.
uint32_t crc32Feed(uint32_t crc, const void* data, size_t len);
inline uint32_t crc32(const void* data, size_t len) { return crc32Feed(0xFFFFFFFFu, data, len) ^ 0xFFFFFFFFu; }

}  // namespace fz
