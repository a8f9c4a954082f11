#include "pinco_format.hpp"

#include <algorithm>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <stdexcept>

namespace fs = std::filesystem;

namespace pinco {

namespace {
const unsigned char kMagic[8] = {0x89, 'P', 'N', 'C', 0x0d, 0x0a, 0x1a, 0x0a};
const unsigned char kVersion = 1;
const size_t kExtra = 16;  // {t0, t1}

template <class T>
bool rd(FILE* f, T& v) { return fread(&v, sizeof(T), 1, f) == 1; }
template <class T>
void wr(FILE* f, const T& v) { fwrite(&v, sizeof(T), 1, f); }

std::string cnk_name(const std::string& folder, size_t i, size_t chunk) {
  return folder + "/variable" + std::to_string(i) + "_" + std::to_string(chunk) + ".cnk";
}
std::string idx_name(const std::string& folder, size_t chunk) {
  return folder + "/index_" + std::to_string(chunk) + ".bin";
}
}  // namespace

PincoFormatDataStore::PincoFormatDataStore(const std::string& folder, double chunk_gb)
    : folder_(folder), chunk_bytes_(chunk_gb * 1024.0 * 1024.0 * 1024.0) {
  fs::create_directories(folder_);
  std::string path = folder_ + "/_register.bin";
  if (fs::exists(path)) {
    reg_ = fopen(path.c_str(), "rb+");
    if (!reg_ || !ReadRegister()) {
      if (reg_) fclose(reg_);
      std::cerr << "ERROR LOADING Register FILE. Creating new Register" << std::endl;
      reg_ = fopen(path.c_str(), "wb+");
      initialised_ = false;
      entries_.clear();
    }
  } else {
    reg_ = fopen(path.c_str(), "wb+");
  }
  if (!reg_) throw std::runtime_error("cannot open " + path);
}

PincoFormatDataStore::~PincoFormatDataStore() {
  for (FILE* f : cnk_) fclose(f);
  if (idx_) fclose(idx_);
  if (reg_) fclose(reg_);
}

bool PincoFormatDataStore::ReadRegister() {
  fseek(reg_, 0, SEEK_SET);
  unsigned char magic[8];
  unsigned char ver = 0;
  if (fread(magic, 1, 8, reg_) != 8 || std::memcmp(magic, kMagic, 8) != 0) return false;
  if (!rd(reg_, ver)) return false;
  int nd = fgetc(reg_);
  if (nd < 0) return false;
  names_.assign(nd, "");
  for (int i = 0; i < nd; ++i) {
    int len = fgetc(reg_);
    if (len < 0) return false;
    names_[i].resize(len);
    if (len && fread(&names_[i][0], 1, len, reg_) != (size_t)len) return false;
  }
  dims_.assign(nd, {});
  for (int i = 0; i < nd; ++i) {
    size_t D = 0;
    if (!rd(reg_, D) || D > 64) return false;  // a corrupt count would allocate without bound
    dims_[i].resize(D);
    for (size_t k = 0; k < D; ++k)
      if (!rd(reg_, dims_[i][k])) return false;
  }
  formats_.assign(nd, 0);
  for (int i = 0; i < nd; ++i) formats_[i] = (unsigned char)fgetc(reg_);
  size_t extra = 0;
  if (!rd(reg_, extra) || !rd(reg_, entries_begin_)) return false;
  if (extra != kExtra) return false;
  fseek(reg_, entries_begin_, SEEK_SET);
  entries_.clear();
  for (;;) {
    long pos = ftell(reg_);
    size_t hdr[4];
    double ex[2];
    if (fread(hdr, sizeof(size_t), 4, reg_) != 4 || fread(ex, sizeof(double), 2, reg_) != 2) break;
    entries_.push_back({hdr[0], hdr[1], hdr[2], hdr[3], ex[0], ex[1], pos});
    chunk_ids_.insert(hdr[1]);
  }
  initialised_ = true;
  return true;
}

void PincoFormatDataStore::InitRegister(const TaskResults& r) {
  fseek(reg_, 0, SEEK_SET);
  fwrite(kMagic, 1, 8, reg_);
  wr(reg_, kVersion);
  names_ = r.NamesOfDatasets();
  if (names_.size() > 255) throw std::runtime_error("too many datasets for the register");
  fputc((int)names_.size(), reg_);
  dims_.clear();
  formats_.clear();
  for (auto& n : names_) {
    if (n.size() > 255) throw std::runtime_error("dataset name too long");
    fputc((int)n.size(), reg_);
    fwrite(n.data(), 1, n.size(), reg_);
  }
  for (size_t i = 0; i < names_.size(); ++i) {
    const auto& d = r.DatasetAt(i);
    size_t D = d.dims.size();
    wr(reg_, D);
    for (size_t k : d.dims) wr(reg_, k);
    dims_.push_back(d.dims);
    formats_.push_back(d.format);
  }
  for (unsigned char f : formats_) fputc(f, reg_);
  size_t extra = kExtra;
  wr(reg_, extra);
  entries_begin_ = ftell(reg_) + (long)sizeof(long);
  wr(reg_, entries_begin_);
  fflush(reg_);
  initialised_ = true;
}

void PincoFormatDataStore::OpenChunk() {
  for (FILE* f : cnk_) fclose(f);
  cnk_.clear();
  if (idx_) fclose(idx_);
  size_t id = chunk_ids_.empty() ? 0 : *chunk_ids_.rbegin() + 1;
  chunk_ids_.insert(id);
  chunk_id_ = (long)id;
  for (size_t i = 0; i < names_.size(); ++i) {
    FILE* f = fopen(cnk_name(folder_, i, id).c_str(), "wb+");
    if (!f) throw std::runtime_error("cannot create chunk file");
    cnk_.push_back(f);
  }
  idx_ = fopen(idx_name(folder_, id).c_str(), "wb+");
  if (!idx_) throw std::runtime_error("cannot create chunk index");
  chunk_written_ = 0;
  chunk_entries_ = 0;
}

void PincoFormatDataStore::StoreData(const TaskResults& r) {
  if (!initialised_) InitRegister(r);
  if (chunk_id_ < 0 || (double)chunk_written_ > chunk_bytes_) OpenChunk();
  std::vector<size_t> rec(1 + 2 * names_.size(), 0);
  rec[0] = r.GetId();
  for (size_t i = 0; i < names_.size(); ++i) {
    const TaskResults::Dataset* d = r.GetDataset(names_[i]);
    FILE* f = cnk_[i];
    fseek(f, 0, SEEK_END);
    size_t off = (size_t)ftell(f);
    size_t nbytes = d ? d->bytes.size() : 0;
    wr(f, nbytes);
    if (nbytes) fwrite(d->bytes.data(), 1, nbytes, f);
    fflush(f);
    rec[1 + 2 * i] = d ? d->frames : 0;
    rec[2 + 2 * i] = off;
    chunk_written_ += nbytes + sizeof(size_t);
  }
  fwrite(rec.data(), sizeof(size_t), rec.size(), idx_);
  fflush(idx_);
  size_t index = chunk_entries_++;

  fseek(reg_, 0, SEEK_END);
  long pos = ftell(reg_);
  size_t hdr[4] = {r.GetId(), (size_t)chunk_id_, index, 0};
  fwrite(hdr, sizeof(size_t), 4, reg_);
  fwrite(r.extraData.data(), sizeof(double), 2, reg_);
  // continuation: point the previous final entry of this id at the new one
  for (auto it = entries_.rbegin(); it != entries_.rend(); ++it) {
    if (it->id == r.GetId() && it->continuation == 0) {
      it->continuation = (size_t)pos;
      fseek(reg_, it->pos + 3 * (long)sizeof(size_t), SEEK_SET);
      wr(reg_, it->continuation);
      break;
    }
  }
  fflush(reg_);
  entries_.push_back({r.GetId(), (size_t)chunk_id_, index, 0, r.extraData[0], r.extraData[1], pos});
}

std::set<size_t> PincoFormatDataStore::UsedIds() const {
  std::set<size_t> s;
  for (auto& e : entries_) s.insert(e.id);
  return s;
}

std::unique_ptr<TaskResults> PincoFormatDataStore::LoadResult(size_t id, bool lastFrameOnly) const {
  std::vector<const Entry*> chain;
  for (auto& e : entries_)
    if (e.id == id) chain.push_back(&e);
  if (chain.empty()) return nullptr;
  if (lastFrameOnly) chain.erase(chain.begin(), chain.end() - 1);
  auto res = std::make_unique<TaskResults>();
  res->SetId(id);
  res->extraData = {chain.front()->t0, chain.back()->t1};
  std::vector<std::vector<unsigned char>> data(names_.size());
  std::vector<size_t> frames(names_.size(), 0);
  for (const Entry* e : chain) {
    FILE* idx = fopen(idx_name(folder_, e->chunk).c_str(), "rb");
    if (!idx) throw std::runtime_error("chunk index " + std::to_string(e->chunk) + " is missing");
    std::vector<size_t> rec(1 + 2 * names_.size());
    fseek(idx, (long)(e->index * rec.size() * sizeof(size_t)), SEEK_SET);
    bool ok = fread(rec.data(), sizeof(size_t), rec.size(), idx) == rec.size();
    fclose(idx);
    if (!ok) throw std::runtime_error("chunk index entry is missing");
    for (size_t i = 0; i < names_.size(); ++i) {
      FILE* f = fopen(cnk_name(folder_, i, e->chunk).c_str(), "rb");
      if (!f) throw std::runtime_error("chunk file " + cnk_name(folder_, i, e->chunk) + " is missing");
      fseek(f, 0, SEEK_END);
      const size_t fsize = (size_t)ftell(f);
      fseek(f, (long)rec[2 + 2 * i], SEEK_SET);
      size_t nbytes = 0;
      if (!rd(f, nbytes) || nbytes > fsize - std::min(fsize, rec[2 + 2 * i] + sizeof(size_t))) {
        fclose(f);
        throw std::runtime_error("corrupt entry in chunk file " + cnk_name(folder_, i, e->chunk));
      }
      size_t nfr = rec[1 + 2 * i];
      std::vector<unsigned char> buf(nbytes);
      if (nbytes && fread(buf.data(), 1, nbytes, f) != nbytes) {
        fclose(f);
        throw std::runtime_error("short read in chunk file");
      }
      fclose(f);
      if (lastFrameOnly && nfr > 1) {
        size_t el = nbytes / nfr;
        buf.erase(buf.begin(), buf.end() - (long)el);
        nfr = 1;
      }
      data[i].insert(data[i].end(), buf.begin(), buf.end());
      frames[i] += nfr;
    }
  }
  for (size_t i = 0; i < names_.size(); ++i)
    res->AddDataset(names_[i], formats_[i], frames[i], dims_[i], data[i].data(), data[i].size());
  return res;
}

void PincoFormatDataStore::SaveFile(const std::string& name, const std::vector<double>& v) const {
  std::ofstream f(folder_ + "/" + name);
  f.precision(17);
  f << "%\tformat:real\tcols:1\trows:" << v.size() << "\n";
  for (double x : v) f << x << "\n";
}

void PincoFormatDataStore::SaveFile(const std::string& name, const std::vector<std::vector<double>>& v) const {
  std::ofstream f(folder_ + "/" + name);
  f.precision(17);
  f << "%\tformat:real\tcols:x\trows:" << v.size() << "\n";
  for (auto& row : v) {
    for (double x : row) f << x << "\t";
    f << "\n";
  }
}

void PincoFormatDataStore::SaveFile(const std::string& name, const std::vector<std::vector<cd>>& v) const {
  std::ofstream f(folder_ + "/" + name);
  f.precision(17);
  f << "%\tformat:complex\tcols:x\trows:" << v.size() << "\n";
  for (auto& row : v) {
    for (cd x : row) f << x.real() << "\t" << x.imag() << "\t";
    f << "\n";
  }
}

}  // namespace pinco
