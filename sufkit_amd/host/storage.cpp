// Width policies, coordinate storage and raw LCP storage.
#include <algorithm>
#include <cstring>
#include <limits>
#include <string>

#include "internal.hpp"

namespace sufkit::detail {

void Throw(ErrorCode code, const std::string& message) { throw Error(code, message); }

std::uint64_t ContentFingerprint(const std::uint8_t* data, std::size_t size) {
  // FNV-1a-64 (reference_data.hpp:46-54)
  std::uint64_t hash = 14695981039346656037ULL;
  for (std::size_t i = 0; i < size; ++i) {
    hash ^= data[i];
    hash *= 1099511628211ULL;
  }
  return hash;
}

// ------------------------------------------------------------- policies ---

SaBackend ResolveSaBackend(SaBackend requested, std::uint64_t, std::uint32_t,
                           bool hip_available) noexcept {
  if (requested != SaBackend::kAutoSelect) return requested;
  // The device backend is the only construction engine of this build; a
  // CPU-only process resolves to it too and reports kUnsupportedBackend.
  (void)hip_available;
  return SaBackend::kHip;
}

CoordinateWidth ResolveSaCoordinateWidth(SaBackend backend, CoordinateWidth requested,
                                         std::uint64_t text_symbols) noexcept {
  if (requested != CoordinateWidth::kAutoSelect) return requested;
  const std::uint64_t limit = backend == SaBackend::kDivsufsort
                                  ? static_cast<std::uint64_t>(std::numeric_limits<std::int32_t>::max())
                                  : static_cast<std::uint64_t>(std::numeric_limits<std::uint32_t>::max());
  return text_symbols <= limit ? CoordinateWidth::kBits32 : CoordinateWidth::kBits64;
}

std::uint64_t MaximumSaConstructionTextSymbols(SaBackend backend, CoordinateWidth width) noexcept {
  if (width == CoordinateWidth::kBits32) {
    if (backend == SaBackend::kCaps || backend == SaBackend::kHip) {
      return std::numeric_limits<std::uint32_t>::max();
    }
    if (backend == SaBackend::kDivsufsort) {
      return static_cast<std::uint64_t>(std::numeric_limits<std::int32_t>::max());
    }
  }
  if (width == CoordinateWidth::kBits64) {
    if (backend == SaBackend::kCaps || backend == SaBackend::kHip) {
      return std::numeric_limits<std::uint64_t>::max();
    }
    if (backend == SaBackend::kDivsufsort) {
      return static_cast<std::uint64_t>(std::numeric_limits<std::int64_t>::max());
    }
  }
  return 0;
}

bool SaConstructionCanRepresent(SaBackend backend, CoordinateWidth width,
                                std::uint64_t text_symbols) noexcept {
  const auto limit = MaximumSaConstructionTextSymbols(backend, width);
  return limit != 0 && text_symbols != 0 && text_symbols <= limit;
}

std::uint64_t MaxCoordinateForWidth(CoordinateStorageWidth width, ErrorCode error_code) {
  switch (width) {
    case CoordinateStorageWidth::kBits32:
      return std::numeric_limits<std::uint32_t>::max();
    case CoordinateStorageWidth::kBits40:
      return kMaxCoordinate40;
    case CoordinateStorageWidth::kBits48:
      return kMaxCoordinate48;
    case CoordinateStorageWidth::kBits64:
      return std::numeric_limits<std::uint64_t>::max();
    case CoordinateStorageWidth::kAutoSelect:
      break;
  }
  Throw(error_code, "storage width must be auto, 32, 40, 48 or 64 bits");
}

CoordinateStorageWidth SelectCoordinateStorageWidth(std::uint64_t symbol_count,
                                                    ErrorCode error_code) {
  if (symbol_count == 0) Throw(error_code, "coordinate domain is empty");
  const std::uint64_t maximum = symbol_count - 1U;
  if (maximum <= std::numeric_limits<std::uint32_t>::max()) return CoordinateStorageWidth::kBits32;
  if (maximum <= kMaxCoordinate40) return CoordinateStorageWidth::kBits40;
  if (maximum <= kMaxCoordinate48) return CoordinateStorageWidth::kBits48;
  return CoordinateStorageWidth::kBits64;
}

CoordinateStorageWidth ResolveCoordinateStorageWidth(CoordinateStorageWidth requested,
                                                     std::uint64_t symbol_count,
                                                     ErrorCode error_code) {
  if (symbol_count == 0) Throw(error_code, "coordinate domain is empty");
  if (requested == CoordinateStorageWidth::kAutoSelect) {
    return SelectCoordinateStorageWidth(symbol_count, error_code);
  }
  if (symbol_count - 1U > MaxCoordinateForWidth(requested, error_code)) {
    Throw(error_code, "text positions do not fit the chosen storage width");
  }
  return requested;
}

CoordinateStorageWidth ResolveProfileStorageWidth(CoordinateStorageWidth requested,
                                                  SaResourceProfile profile,
                                                  std::uint64_t symbol_count) {
  if (profile != SaResourceProfile::kFast && profile != SaResourceProfile::kLowMemory) {
    Throw(ErrorCode::kInvalidInput, "resource profile must be Fast or Low-memory");
  }
  if (requested != CoordinateStorageWidth::kAutoSelect) {
    return ResolveCoordinateStorageWidth(requested, symbol_count);
  }
  const auto narrowest = SelectCoordinateStorageWidth(symbol_count);
  if (profile == SaResourceProfile::kFast && narrowest != CoordinateStorageWidth::kBits32) {
    return CoordinateStorageWidth::kBits64;
  }
  return narrowest;
}

CoordinateStorageWidth ResolveAuxiliaryStorageWidth(CoordinateStorageWidth preferred,
                                                    std::uint64_t symbol_count) {
  if (symbol_count - 1U <= MaxCoordinateForWidth(preferred)) return preferred;
  return SelectCoordinateStorageWidth(symbol_count);
}

// ----------------------------------------------------- coordinate storage ---

namespace {

template <class Values>
void CheckRange(const Values& values, std::uint64_t symbol_count, ErrorCode error_code,
                const char* label) {
  for (const auto v : values) {
    if (static_cast<std::uint64_t>(v) >= symbol_count) {
      Throw(error_code, std::string(label ? label : "coordinate") +
                            " contains a value outside its domain");
    }
  }
}

template <class Values>
CoordinateStorage::Storage Pack(Values&& values, CoordinateStorageWidth width) {
  using T = typename std::decay_t<Values>::value_type;
  if (width == CoordinateStorageWidth::kBits32) {
    if constexpr (std::is_same_v<T, std::uint32_t>) {
      return std::move(values);
    } else {
      U32Values out(values.size());
      for (std::size_t i = 0; i < values.size(); ++i) out[i] = static_cast<std::uint32_t>(values[i]);
      return out;
    }
  }
  if (width == CoordinateStorageWidth::kBits64) {
    if constexpr (std::is_same_v<T, std::uint64_t>) {
      return std::move(values);
    } else {
      return U64Values(values.begin(), values.end());
    }
  }
  if (width == CoordinateStorageWidth::kBits40) {
    Coordinate40 out;
    out.low.resize(values.size());
    out.high.resize(values.size());
    for (std::size_t i = 0; i < values.size(); ++i) {
      const std::uint64_t v = values[i];
      out.low[i] = static_cast<std::uint32_t>(v);
      out.high[i] = static_cast<std::uint8_t>(v >> 32);
    }
    return out;
  }
  Coordinate48 out;
  out.low.resize(values.size());
  out.high.resize(values.size());
  for (std::size_t i = 0; i < values.size(); ++i) {
    const std::uint64_t v = values[i];
    out.low[i] = static_cast<std::uint32_t>(v);
    out.high[i] = static_cast<std::uint16_t>(v >> 32);
  }
  return out;
}

}  // namespace

CoordinateStorage CoordinateStorage::FromUInt32(U32Values&& values,
                                                CoordinateStorageWidth width,
                                                std::uint64_t symbol_count, ErrorCode error_code,
                                                const char* label) {
  width = ResolveCoordinateStorageWidth(width, symbol_count, error_code);
  CheckRange(values, symbol_count, error_code, label);
  CoordinateStorage s;
  s.storage_ = Pack(std::move(values), width);
  s.width_ = width;
  return s;
}

CoordinateStorage CoordinateStorage::FromUInt64(U64Values&& values,
                                                CoordinateStorageWidth width,
                                                std::uint64_t symbol_count, ErrorCode error_code,
                                                const char* label) {
  width = ResolveCoordinateStorageWidth(width, symbol_count, error_code);
  CheckRange(values, symbol_count, error_code, label);
  CoordinateStorage s;
  s.storage_ = Pack(std::move(values), width);
  s.width_ = width;
  return s;
}

CoordinateStorage CoordinateStorage::FromPacked40(Coordinate40&& values, std::uint64_t symbol_count,
                                                  ErrorCode error_code, const char* label) {
  if (values.low.size() != values.high.size()) Throw(error_code, "split40 planes differ in length");
  (void)ResolveCoordinateStorageWidth(CoordinateStorageWidth::kBits40, symbol_count, error_code);
  for (std::size_t i = 0; i < values.low.size(); ++i) {
    const std::uint64_t v = (static_cast<std::uint64_t>(values.high[i]) << 32) | values.low[i];
    if (v >= symbol_count) Throw(error_code, std::string(label ? label : "coordinate") + " out of range");
  }
  CoordinateStorage s;
  s.storage_ = std::move(values);
  s.width_ = CoordinateStorageWidth::kBits40;
  return s;
}

CoordinateStorage CoordinateStorage::FromPacked48(Coordinate48&& values, std::uint64_t symbol_count,
                                                  ErrorCode error_code, const char* label) {
  if (values.low.size() != values.high.size()) Throw(error_code, "split48 planes differ in length");
  (void)ResolveCoordinateStorageWidth(CoordinateStorageWidth::kBits48, symbol_count, error_code);
  for (std::size_t i = 0; i < values.low.size(); ++i) {
    const std::uint64_t v = (static_cast<std::uint64_t>(values.high[i]) << 32) | values.low[i];
    if (v >= symbol_count) Throw(error_code, std::string(label ? label : "coordinate") + " out of range");
  }
  CoordinateStorage s;
  s.storage_ = std::move(values);
  s.width_ = CoordinateStorageWidth::kBits48;
  return s;
}

std::uint64_t CoordinateStorage::Size() const noexcept {
  switch (storage_.index()) {
    case 1: return std::get<1>(storage_).size();
    case 2: return std::get<2>(storage_).size();
    case 3: return std::get<3>(storage_).low.size();
    case 4: return std::get<4>(storage_).low.size();
    default: return 0;
  }
}

std::uint64_t CoordinateStorage::Bytes() const noexcept {
  switch (width_) {
    case CoordinateStorageWidth::kBits32: return Size() * 4;
    case CoordinateStorageWidth::kBits40: return Size() * 5;
    case CoordinateStorageWidth::kBits48: return Size() * 6;
    case CoordinateStorageWidth::kBits64: return Size() * 8;
    default: return 0;
  }
}

std::uint64_t CoordinateStorage::At(std::uint64_t i) const noexcept {
  switch (storage_.index()) {
    case 1: return std::get<1>(storage_)[i];
    case 2: return std::get<2>(storage_)[i];
    case 3: {
      const auto& v = std::get<3>(storage_);
      return (static_cast<std::uint64_t>(v.high[i]) << 32) | v.low[i];
    }
    case 4: {
      const auto& v = std::get<4>(storage_);
      return (static_cast<std::uint64_t>(v.high[i]) << 32) | v.low[i];
    }
    default: return 0;
  }
}

// ------------------------------------------------------------ LCP ---

LcpStorage LcpStorage::FromRaw32(U32Values&& values, std::uint32_t sampling_rate,
                                 ErrorCode error_code) {
  if (sampling_rate == 0) Throw(error_code, "LCP storage: sampling rate must be positive");
  if (!values.empty() && values.front() != 0) Throw(error_code, "LCP storage: LCP[0] must be zero");
  LcpStorage s;
  s.raw32_ = std::move(values);
  s.width_ = 32;
  s.sampling_rate_ = sampling_rate;
  return s;
}

LcpStorage LcpStorage::FromRaw64(U64Values&& values, std::uint32_t sampling_rate,
                                 ErrorCode error_code) {
  if (sampling_rate == 0) Throw(error_code, "LCP storage: sampling rate must be positive");
  if (!values.empty() && values.front() != 0) Throw(error_code, "LCP storage: LCP[0] must be zero");
  LcpStorage s;
  s.raw64_ = std::move(values);
  s.width_ = 64;
  s.sampling_rate_ = sampling_rate;
  return s;
}

LcpStorage LcpStorage::FromByteCoded(std::vector<std::uint8_t>&& primary, std::vector<std::uint64_t>&& positions,
                                     std::vector<std::uint64_t>&& values, std::uint32_t sampling_rate,
                                     std::uint64_t text_symbols, std::uint8_t coordinate_width,
                                     ErrorCode error_code) {
  if (sampling_rate == 0) Throw(error_code, "LCP storage: sampling rate must be positive");
  if (coordinate_width != 32 && coordinate_width != 64) Throw(error_code, "LCP storage: invalid coordinate width");
  if (positions.size() != values.size()) {
    Throw(error_code, "LCP storage: anchor positions and values differ in number");
  }
  const std::uint64_t max_coordinate =
      coordinate_width == 32 ? std::numeric_limits<std::uint32_t>::max() : std::numeric_limits<std::uint64_t>::max();
  if (text_symbols != 0 && text_symbols - 1 > max_coordinate) {
    Throw(error_code, "LCP storage: anchor width too narrow for the text");
  }
  const std::uint64_t rows = text_symbols == 0 ? 0 : 1 + (text_symbols - 1) / sampling_rate;
  if (primary.size() != rows) Throw(error_code, "LCP storage: row count does not follow from the sampling rate");
  if (positions.size() > (primary.empty() ? 0 : primary.size() - 1)) {
    Throw(error_code, "LCP storage: more anchors than rows");
  }
  if (!primary.empty() && primary.front() != 0) Throw(error_code, "LCP storage: LCP[0] must be zero");
  for (std::size_t i = 0; i < positions.size(); ++i) {
    const std::uint64_t p = positions[i], v = values[i];
    if (p >= text_symbols) Throw(error_code, "LCP storage: an anchor position lies past the text");
    if (p % sampling_rate != 0) Throw(error_code, "LCP storage: an anchor position is not a multiple of the sampling rate");
    if (v < kByteLimit) Throw(error_code, "LCP storage: an anchor holds a value below 255");
    if (v > max_coordinate) Throw(error_code, "LCP storage: an anchor value does not fit the anchor width");
    if (i == 0) continue;
    if (p <= positions[i - 1]) Throw(error_code, "LCP storage: anchor positions are not in increasing order");
    const std::uint64_t delta = p - positions[i - 1];
    if (delta == sampling_rate && values[i - 1] >= delta && values[i - 1] - delta == v) {
      Throw(error_code, "LCP storage: an anchor sits where the previous run continues");
    }
  }
  LcpStorage s;
  s.byte_coded_ = true;
  s.width_ = coordinate_width;
  s.sampling_rate_ = sampling_rate;
  s.primary_ = std::move(primary);
  s.anchor_positions_ = std::move(positions);
  s.anchor_values_ = std::move(values);
  if (text_symbols != 0) {
    // guide[b] = first anchor at or after position b * 4096, b = 0..blocks.
    const std::uint64_t blocks = 1 + ((text_symbols - 1) >> kGuideBlockShift);
    s.guide_.resize(blocks + 1);
    std::size_t a = 0;
    for (std::uint64_t b = 0; b <= blocks; ++b) {
      const std::uint64_t boundary = b << kGuideBlockShift;
      while (a < s.anchor_positions_.size() && s.anchor_positions_[a] < boundary) ++a;
      s.guide_[b] = a;
    }
  }
  return s;
}

std::uint64_t LcpStorage::Size() const noexcept {
  if (byte_coded_) return primary_.size();
  return width_ == 32 ? raw32_.size() : (width_ == 64 ? raw64_.size() : 0);
}

std::uint64_t LcpStorage::Bytes() const noexcept {
  if (!byte_coded_) return Size() * (width_ / 8u);
  return primary_.size() + (2 * anchor_positions_.size() + guide_.size()) * (width_ / 8u);
}

std::uint64_t LcpStorage::At(std::uint64_t row) const noexcept {
  return width_ == 32 ? raw32_[row] : raw64_[row];
}

std::uint64_t LcpStorage::Decode(std::uint64_t row, std::uint64_t suffix_position, ErrorCode error_code) const {
  if (!byte_coded_) return At(row);
  if (row >= primary_.size()) Throw(error_code, "LCP storage: row is out of range");
  const std::uint8_t primary = primary_[row];
  if (primary < kByteLimit) return primary;
  if (suffix_position % sampling_rate_ != 0) Throw(error_code, "LCP storage: suffix position is not aligned");
  if (anchor_positions_.empty()) Throw(error_code, "LCP storage: a 255 row has no anchor to decode it");
  const std::uint64_t block = suffix_position >> kGuideBlockShift;
  if (block + 1 >= guide_.size()) Throw(error_code, "LCP storage: suffix position lies beyond the anchor guide");
  const auto first = anchor_positions_.begin() + static_cast<std::ptrdiff_t>(guide_[block]);
  const auto last = anchor_positions_.begin() + static_cast<std::ptrdiff_t>(guide_[block + 1]);
  const auto found = std::upper_bound(first, last, suffix_position);
  std::size_t anchor;
  if (found != first) {
    anchor = static_cast<std::size_t>(found - anchor_positions_.begin()) - 1;
  } else {
    if (guide_[block] == 0) Throw(error_code, "LCP storage: a 255 row comes before the first anchor");
    anchor = static_cast<std::size_t>(guide_[block]) - 1;
  }
  const std::uint64_t at = anchor_positions_[anchor], value = anchor_values_[anchor];
  const std::uint64_t delta = suffix_position - at;
  if (at > suffix_position || delta % sampling_rate_ != 0 || value < delta || value - delta < kByteLimit) {
    Throw(error_code, "LCP storage: no anchor run reaches this suffix position");
  }
  return value - delta;
}

void LcpStorage::Validate(const CoordinateStorage& sa, std::uint64_t text_symbols,
                          ErrorCode error_code) const {
  if (Size() != sa.Size()) Throw(error_code, "LCP storage: LCP rows and SA rows differ in number");
  if (Size() == 0) return;
  if (text_symbols == 0 || Size() != 1 + (text_symbols - 1) / sampling_rate_) {
    Throw(error_code, "LCP storage: rows do not follow from the sampling rate");
  }
  if (Decode(0, sa.At(0), error_code) != 0) Throw(error_code, "LCP storage: LCP[0] must be zero");
  for (std::uint64_t row = 1; row < Size(); ++row) {
    const std::uint64_t suffix = sa.At(row), previous = sa.At(row - 1);
    if (suffix >= text_symbols || previous >= text_symbols) {
      Throw(error_code, "LCP storage: an SA value lies past the text");
    }
    if (Decode(row, suffix, error_code) > std::min(text_symbols - suffix, text_symbols - previous)) {
      Throw(error_code, "LCP storage: an LCP runs past the end of its suffixes");
    }
  }
}

}  // namespace sufkit::detail
