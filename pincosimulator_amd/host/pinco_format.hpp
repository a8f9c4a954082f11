// pinco_format.hpp -- the reference's binary output format ("PincoFormat").
//
// Layout as written by src/Base/FileFormats/PincoFormat/ChunkRegister.cpp:152-233,
// 312-381 and ChunkFileSet.cpp:58-160, in this project's own code:
//   _register.bin   magic 89 'P' 'N' 'C' 0D 0A 1A 0A, version byte 1,
//                   1-byte dataset count, per dataset a 1-byte name length and
//                   the name, per dataset size_t D and D size_t frame dims,
//                   per dataset a 1-byte format (22 complex128, 11 float64),
//                   size_t extra bytes per entry (16), size_t offset of the
//                   first entry, then entries of 4 size_t {id, chunk id,
//                   index in the chunk, continuation offset} + {t0, t1}.
//   variable<i>_<chunk>.cnk   per entry: size_t byte count + raw bytes.
//   index_<chunk>.bin         per entry: size_t id + per dataset
//                             {frames, byte offset of the entry in variable<i>}.
// A trajectory continued later gets a second entry; the previous entry's
// continuation offset then points at the new entry's register position.
#ifndef PINCO_FORMAT_HPP
#define PINCO_FORMAT_HPP

#include <cstdio>
#include <map>
#include <memory>
#include <set>
#include <string>
#include <vector>

#include "noisy_matrix.hpp"
#include "plugin.hpp"

namespace pinco {

class PincoFormatDataStore {
 public:
  explicit PincoFormatDataStore(const std::string& folder, double chunk_gb = 1.5);
  ~PincoFormatDataStore();

  void StoreData(const TaskResults& r);
  std::set<size_t> UsedIds() const;
  std::unique_ptr<TaskResults> LoadResult(size_t id, bool lastFrameOnly) const;
  std::unique_ptr<TaskResults> LoadEndFrame(size_t id) const { return LoadResult(id, true); }

  // text side files (DataStore.cpp:70-135)
  void SaveFile(const std::string& name, const std::vector<double>& v) const;
  void SaveFile(const std::string& name, const std::vector<std::vector<double>>& v) const;
  void SaveFile(const std::string& name, const std::vector<std::vector<cd>>& v) const;
  const std::string& Folder() const { return folder_; }

 private:
  struct Entry {
    size_t id, chunk, index, continuation;
    double t0, t1;
    long pos;  // register position of this entry
  };
  bool ReadRegister();
  void InitRegister(const TaskResults& r);
  void OpenChunk();
  std::string folder_;
  double chunk_bytes_;
  FILE* reg_ = nullptr;
  bool initialised_ = false;
  std::vector<std::string> names_;
  std::vector<std::vector<size_t>> dims_;
  std::vector<unsigned char> formats_;
  long entries_begin_ = 0;
  std::vector<Entry> entries_;
  std::set<size_t> chunk_ids_;
  // current write chunk
  long chunk_id_ = -1;
  std::vector<FILE*> cnk_;
  FILE* idx_ = nullptr;
  size_t chunk_written_ = 0, chunk_entries_ = 0;
};

}  // namespace pinco

#endif
