#include "settings.hpp"

#include <cmath>
#include <ctime>
#include <filesystem>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>

#include "noisy_matrix.hpp"

namespace fs = std::filesystem;

namespace pinco {

namespace {

std::string trim(const std::string& s) {
  size_t a = s.find_first_not_of(" \t\r\n");
  if (a == std::string::npos) return "";
  size_t b = s.find_last_not_of(" \t\r\n");
  return s.substr(a, b - a + 1);
}

// strips a # comment that is not inside a string
std::string strip_comment(const std::string& line) {
  bool in_s = false;
  char q = 0;
  for (size_t i = 0; i < line.size(); ++i) {
    char c = line[i];
    if (in_s) {
      if (c == q) in_s = false;
    } else if (c == '"' || c == '\'') {
      in_s = true;
      q = c;
    } else if (c == '#') {
      return line.substr(0, i);
    }
  }
  return line;
}

std::string unquote_key(std::string k) {
  k = trim(k);
  if (k.size() >= 2 && (k.front() == '"' || k.front() == '\'') && k.back() == k.front())
    return k.substr(1, k.size() - 2);
  return k;
}

void parse_into(std::map<std::string, Value>& cfg, const std::string& text) {
  std::istringstream in(text);
  std::string line, table;
  int lineno = 0;
  while (std::getline(in, line)) {
    ++lineno;
    line = trim(strip_comment(line));
    if (line.empty()) continue;
    if (line.front() == '[') {
      size_t e = line.find(']');
      if (e == std::string::npos) throw std::runtime_error("bad table header at line " + std::to_string(lineno));
      table = unquote_key(line.substr(1, e - 1));
      continue;
    }
    // key = value; the key may be quoted and contain '='
    size_t eq = std::string::npos;
    bool in_s = false;
    char q = 0;
    for (size_t i = 0; i < line.size(); ++i) {
      char c = line[i];
      if (in_s) {
        if (c == q) in_s = false;
      } else if (c == '"' || c == '\'') {
        in_s = true;
        q = c;
      } else if (c == '=') {
        eq = i;
        break;
      }
    }
    if (eq == std::string::npos) throw std::runtime_error("expected key = value at line " + std::to_string(lineno));
    std::string key = unquote_key(line.substr(0, eq));
    Value v;
    if (!ParseValue(trim(line.substr(eq + 1)), v))
      throw std::runtime_error("cannot parse the value of '" + key + "' at line " + std::to_string(lineno));
    cfg[table.empty() ? key : table + "." + key] = v;
  }
}

}  // namespace

bool ParseValue(const std::string& t0, Value& out) {
  std::string t = trim(t0);
  if (t.empty()) return false;
  if (t.front() == '"' || t.front() == '\'') {
    if (t.size() < 2 || t.back() != t.front()) return false;
    out.kind = Value::String;
    out.s = t.substr(1, t.size() - 2);
    return true;
  }
  if (t == "true" || t == "false") {
    out.kind = Value::Bool;
    out.b = (t == "true");
    return true;
  }
  if (t.front() == '[') {
    if (t.back() != ']') return false;
    out.kind = Value::Array;
    out.arr.clear();
    std::string body = t.substr(1, t.size() - 2);
    std::stringstream ss(body);
    std::string item;
    while (std::getline(ss, item, ',')) {
      item = trim(item);
      if (item.empty()) continue;
      Value e;
      if (!ParseValue(item, e) || (e.kind != Value::Int && e.kind != Value::Float)) return false;
      out.arr.push_back(e.number());
    }
    return true;
  }
  std::string num;
  for (char c : t)
    if (c != '_') num += c;
  bool is_float = num.find_first_of(".eE") != std::string::npos || num == "inf" || num == "+inf" ||
                  num == "-inf" || num == "nan" || num == "+nan" || num == "-nan";
  try {
    size_t pos = 0;
    if (is_float) {
      out.f = std::stod(num, &pos);
      out.kind = Value::Float;
    } else {
      out.i = std::stoll(num, &pos, 10);
      out.kind = Value::Int;
    }
    return pos == num.size();
  } catch (...) {
    return false;
  }
}

Settings Settings::FromString(const std::string& text, const std::string& basePath) {
  Settings s;
  parse_into(s.config_, text);
  s.basePath_ = basePath.empty() ? "." : basePath;
  s.inputFolder_ = s.basePath_;
  s.outputFolder_ = s.basePath_;
  return s;
}

void Settings::Merge(const std::string& text) { parse_into(config_, text); }

// -i/--input <file or folder> [-o/--output <folder>] [-a/--append <folder>]
// [--dry-run] [key value]...  A folder input holds a previous run's
// _parameters.ini (a subsequent run, Settings.cpp:68-79).
Settings::Settings(int argc, char** argv) {
  std::string input, output;
  std::vector<std::string> extras;
  for (int a = 1; a < argc; ++a) {
    std::string s = argv[a];
    auto need = [&](const char* what) {
      if (a + 1 >= argc) throw std::runtime_error(std::string(what) + " needs an argument");
      return std::string(argv[++a]);
    };
    if (s == "-i" || s == "--input" || s == "input") input = need("-i");
    else if (s == "-o" || s == "--output" || s == "-a" || s == "--append") {
      output = need("-o");
      outputSet_ = true;
    } else if (s == "--dry-run") dryRun_ = true;
    else extras.push_back(s);
  }
  if (input.empty()) throw std::runtime_error("usage: twmc_sim -i input.ini [-o output] [key value ...]");
  fs::path in(input);
  if (fs::is_directory(in)) {
    if (!outputSet_) output = in.string();
    outputSet_ = true;
    in = in / "_parameters.ini";
    subsequentRun_ = true;
  }
  if (!fs::exists(in)) throw std::runtime_error("input file " + in.string() + " does not exist");
  inputPath_ = in.string();
  basePath_ = in.parent_path().string();
  if (basePath_.empty()) basePath_ = ".";
  inputFolder_ = basePath_;
  outputFolder_ = outputSet_ ? output : basePath_;
  std::ifstream f(inputPath_);
  std::stringstream buf;
  buf << f.rdbuf();
  parse_into(config_, buf.str());
  if (extras.size() % 2 != 0) throw std::runtime_error("overridden parameters must have an argument");
  std::string merged;
  for (size_t k = 0; k < extras.size(); k += 2) merged += extras[k] + " = " + extras[k + 1] + "\n";
  if (!merged.empty()) parse_into(config_, merged);
}

const Value* Settings::find(const std::string& key) const {
  auto it = config_.find(key);
  return it == config_.end() ? nullptr : &it->second;
}

double Settings::getDouble(const std::string& key, double def) const {
  const Value* v = find(key);
  if (!v || (v->kind != Value::Int && v->kind != Value::Float && v->kind != Value::Bool)) return def;
  return v->number();
}

int64_t Settings::getInt(const std::string& key, int64_t def) const {
  const Value* v = find(key);
  if (!v) return def;
  if (v->kind == Value::Int) return v->i;
  if (v->kind == Value::Float) return (int64_t)v->f;
  if (v->kind == Value::Bool) return v->b;
  return def;
}

bool Settings::getBool(const std::string& key, bool def) const {
  const Value* v = find(key);
  if (!v) return def;
  if (v->kind == Value::Bool) return v->b;
  if (v->kind == Value::Int) return v->i != 0;
  return def;
}

std::string Settings::getString(const std::string& key, const std::string& def) const {
  const Value* v = find(key);
  return (v && v->kind == Value::String) ? v->s : def;
}

std::complex<double> Settings::getComplex(const std::string& key, std::complex<double> def) const {
  const Value* v = find(key);
  if (v && (v->kind == Value::Int || v->kind == Value::Float)) return {v->number(), 0.0};
  const Value* r = find(key + ".real");
  const Value* i = find(key + ".imag");
  if (!r && !i) return def;
  return {r ? r->number() : def.real(), i ? i->number() : def.imag()};
}

std::vector<double> Settings::getArray(const std::string& key) const {
  const Value* v = find(key);
  if (v && v->kind == Value::Array) return v->arr;
  return {};
}

std::vector<std::string> Settings::keys() const {
  std::vector<std::string> k;
  for (auto& kv : config_) k.push_back(kv.first);
  return k;
}

std::string Settings::GetRootFolder() const {
  if (subsequentRun_) return basePath_;
  std::string r = getString("rootFolder", inputFolder_);
  return r.empty() ? "." : r;
}

uint64_t Settings::GlobalSeed() const {
  uint64_t s = (uint64_t)getInt("SEED", 0);
  if (getBool("seed_add_time", false)) s += (uint64_t)time(nullptr);
  return s;
}

std::unique_ptr<NoisyMatrix> Settings::GetMatrix(const std::string& path, size_t nx, size_t ny,
                                                 size_t cellSz) const {
  auto m = std::make_unique<NoisyMatrix>(nx, ny, cellSz);
  fs::path root = GetRootFolder();
  auto file_values = [&](const std::string& rel, bool& ok) {
    fs::path p = root / rel;
    ok = fs::exists(p);
    return ok ? ReadCharFile(p.string()) : std::vector<double>();
  };
  bool ok = true;
  if (contains(path + ".data")) {
    const Value* v = find(path + ".data");
    if (v->kind == Value::Array) ok = m->SetValue(v->arr);
    else m->SetValue(cd(v->number(), 0.0));
  } else if (contains(path + ".file")) {
    auto vals = file_values(getString(path + ".file"), ok);
    if (ok) ok = m->SetValue(vals);
    else std::cerr << "FILE NOT FOUND! " << (root / getString(path + ".file")) << std::endl;
  } else if (contains(path + ".t") || contains(path + "_t")) {
    std::string rel = contains(path + ".t") ? getString(path + ".t") : getString(path + "_t");
    fs::path p = root / rel;
    if (fs::exists(p)) m->SetTemporalValue(ReadTemporalCharFile(p.string()));
    else {
      ok = false;
      std::cerr << "FILE NOT FOUND! " << p << std::endl;
    }
  } else if (contains(path + ".real")) {
    m->SetValue(getComplex(path, cd(0.0, 0.0)));
  } else if (const Value* v = find(path)) {
    if (v->kind == Value::String) {
      auto vals = file_values(v->s, ok);
      if (ok) ok = m->SetValue(vals);
      else std::cerr << "FILE NOT FOUND! " << (root / v->s) << std::endl;
    } else if (v->kind == Value::Array) {
      ok = m->SetValue(v->arr);
    } else {
      m->SetValue(cd(v->number(), 0.0));
    }
  } else {
    ok = false;
  }
  if (!ok)
    throw std::runtime_error("matrix parameter '" + path + "' is missing or unreadable");

  // noise: X.Noise_Type / X.Noise_Val_0, or the X_Noise_Type / X_Noise_Val_0 form
  for (const char* sep : {".", "_"}) {
    std::string tkey = path + sep + "Noise_Type";
    if (!contains(tkey)) continue;
    m->SetNoiseType(getString(tkey, "None"));
    for (size_t i = 0; i < m->GetNoiseVariables(); ++i) {
      std::string vkey = path + sep + "Noise_Val_" + std::to_string(i);
      const Value* v = find(vkey);
      bool set = false;
      if (v && v->kind == Value::String) {
        bool fok = false;
        auto vals = file_values(v->s, fok);
        set = fok && m->SetNoiseVal(i, vals);
      } else if (v || contains(vkey + ".real")) {
        m->SetNoiseVal(i, getComplex(vkey, cd(0.0, 0.0)));
        set = true;
      }
      if (!set) throw std::runtime_error("Variable " + vkey + " was not set, but was necessary");
    }
    break;
  }
  return m;
}

std::vector<double> ReadCharFile(const std::string& path) {
  std::ifstream f(path);
  std::vector<double> v;
  std::string tok;
  while (f >> tok) {
    if (tok == "*" || tok == "nan" || tok == "NaN" || tok == "NAN") v.push_back(NAN);
    else v.push_back(std::stod(tok));
  }
  return v;
}

// blocks of "time" line followed by value lines, separated by an empty line
// or "-" (Settings.cpp:530-560); '*' and nan mark unknown values
std::map<double, std::vector<double>> ReadTemporalCharFile(const std::string& path) {
  std::map<double, std::vector<double>> out;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    if (trim(line).empty()) continue;
    double t = std::stod(trim(line));
    std::vector<double> vals;
    while (std::getline(f, line)) {
      std::string tl = trim(line);
      if (tl.empty() || tl == "-") break;
      std::istringstream iss(tl);
      std::string tok;
      while (iss >> tok) {
        if (tok == "*" || tok == "nan" || tok == "NaN" || tok == "NAN") vals.push_back(NAN);
        else vals.push_back(std::stod(tok));
      }
    }
    out[t] = vals;
  }
  return out;
}

}  // namespace pinco
