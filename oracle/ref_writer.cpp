// C entry points over the compiled reference library: build with the
// reference's own SuffixArray::Build (divsufsort backend) and Save a .sufidx
// 1.4 file, or Load one and report its SA/ISA/LCP.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target ref-full)
// against the reference sources where they lie; nothing of the reference is
// copied into this repository.  Used by tests/golden/make_golden.py to write
// golden .sufidx files, so the writer of this repository can be compared
// with the reference's byte for byte.
#include <cstdint>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include <sufkit/sufkit.hpp>

extern "C" {

// acceleration: SaAcceleration value; coordinate_width 0/32/64; storage 0/32/40/48/64.
int sfk_ref_write_sufidx(const char* const* names, const char* const* sequences,
                         const std::uint64_t* lengths, std::uint64_t count, const char* path,
                         int coordinate_width, int storage_width, int acceleration,
                         char* error, std::size_t capacity) {
  try {
    std::vector<sufkit::SequenceRecord> records(count);
    for (std::uint64_t i = 0; i < count; ++i) {
      records[i].name = names[i];
      records[i].sequence.assign(sequences[i], lengths[i]);
    }
    const auto reference = sufkit::GenomeReference::FromRecords(std::move(records));
    sufkit::SuffixArrayBuildOptions options;
    options.backend = sufkit::SaBackend::kDivsufsort;
    options.coordinate_width = static_cast<sufkit::CoordinateWidth>(coordinate_width);
    options.storage_width = static_cast<sufkit::CoordinateStorageWidth>(storage_width);
    options.acceleration = static_cast<sufkit::SaAcceleration>(acceleration);
    const auto index = sufkit::SuffixArray::Build(reference, options);
    sufkit::SaveOptions save;
    save.overwrite = true;
    index.Save(path, save);
    return 0;
  } catch (const std::exception& e) {
    if (error != nullptr && capacity > 0) {
      std::strncpy(error, e.what(), capacity - 1);
      error[capacity - 1] = '\0';
    }
    return -1;
  }
}

}  // extern "C"
