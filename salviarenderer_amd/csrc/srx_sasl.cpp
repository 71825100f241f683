// SASL pixel-shader compiler (see srx_sasl.h): a lexer, a recursive-descent parser that emits
// code statement by statement, and a register allocator over the quad interpreter's 32 registers.
#include "srx_sasl.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <memory>

namespace srx {
namespace {

// ---------------------------------------------------------------- lexer
enum tok_kind { TK_EOF, TK_ID, TK_NUM, TK_PUNCT };
struct token {
  tok_kind k;
  std::string s;
  double num;
  bool is_float;
  int line;
};

struct compile_error {
  int status;
  std::string msg;
};
[[noreturn]] void raise(int status, int line, const std::string& m) {
  throw compile_error{status, "line " + std::to_string(line) + ": " + m};
}

std::vector<token> lex(const char* p, size_t n) {
  std::vector<token> out;
  size_t i = 0;
  int line = 1;
  static const char* const multi[] = {"<=", ">=", "==", "!=", "&&", "||", "+=", "-=", "*=", "/=", "%=",
                                      "++", "--", "<<", ">>", "&=", "|=", "^="};
  while (i < n) {
    const char c = p[i];
    if (c == '\n') {
      ++line;
      ++i;
      continue;
    }
    if (c == ' ' || c == '\t' || c == '\r' || c == '\f' || c == '\v') {
      ++i;
      continue;
    }
    if (c == '/' && i + 1 < n && p[i + 1] == '/') {
      while (i < n && p[i] != '\n') ++i;
      continue;
    }
    if (c == '/' && i + 1 < n && p[i + 1] == '*') {
      i += 2;
      while (i + 1 < n && !(p[i] == '*' && p[i + 1] == '/')) {
        if (p[i] == '\n') ++line;
        ++i;
      }
      if (i + 1 >= n) raise(SASL_ERROR, line, "unterminated comment");
      i += 2;
      continue;
    }
    if (c == '#') raise(SASL_UNSUPPORTED, line, "preprocessor directives are not supported");
    if (std::isalpha(static_cast<unsigned char>(c)) || c == '_') {
      size_t j = i;
      while (j < n && (std::isalnum(static_cast<unsigned char>(p[j])) || p[j] == '_')) ++j;
      out.push_back({TK_ID, std::string(p + i, j - i), 0.0, false, line});
      i = j;
      continue;
    }
    if (std::isdigit(static_cast<unsigned char>(c)) || (c == '.' && i + 1 < n && std::isdigit(static_cast<unsigned char>(p[i + 1])))) {
      size_t j = i;
      bool is_float = false;
      if (c == '0' && j + 1 < n && (p[j + 1] == 'x' || p[j + 1] == 'X')) {
        j += 2;
        while (j < n && std::isxdigit(static_cast<unsigned char>(p[j]))) ++j;
        const std::string s(p + i, j - i);
        out.push_back({TK_NUM, s, static_cast<double>(std::strtoull(s.c_str(), nullptr, 16)), false, line});
        while (j < n && (p[j] == 'u' || p[j] == 'U' || p[j] == 'l' || p[j] == 'L')) ++j;
        i = j;
        continue;
      }
      while (j < n && std::isdigit(static_cast<unsigned char>(p[j]))) ++j;
      if (j < n && p[j] == '.') {
        is_float = true;
        ++j;
        while (j < n && std::isdigit(static_cast<unsigned char>(p[j]))) ++j;
      }
      if (j < n && (p[j] == 'e' || p[j] == 'E')) {
        size_t k = j + 1;
        if (k < n && (p[k] == '+' || p[k] == '-')) ++k;
        if (k < n && std::isdigit(static_cast<unsigned char>(p[k]))) {
          is_float = true;
          j = k;
          while (j < n && std::isdigit(static_cast<unsigned char>(p[j]))) ++j;
        }
      }
      const std::string s(p + i, j - i);
      if (j < n && (p[j] == 'f' || p[j] == 'F' || p[j] == 'h' || p[j] == 'H')) {
        is_float = true;
        ++j;
      } else {
        while (j < n && (p[j] == 'u' || p[j] == 'U' || p[j] == 'l' || p[j] == 'L')) ++j;
      }
      out.push_back({TK_NUM, s, std::strtod(s.c_str(), nullptr), is_float, line});
      i = j;
      continue;
    }
    bool matched = false;
    for (const char* m : multi) {
      if (i + 1 < n && p[i] == m[0] && p[i + 1] == m[1]) {
        out.push_back({TK_PUNCT, m, 0.0, false, line});
        i += 2;
        matched = true;
        break;
      }
    }
    if (matched) continue;
    if (std::strchr("+-*/%=<>!(){}[],;:.?&|^~", c)) {
      out.push_back({TK_PUNCT, std::string(1, c), 0.0, false, line});
      ++i;
      continue;
    }
    raise(SASL_ERROR, line, std::string("unexpected character '") + c + "'");
  }
  out.push_back({TK_EOF, "", 0.0, false, line});
  return out;
}

// ---------------------------------------------------------------- types
enum base_t { BT_VOID, BT_BOOL, BT_INT, BT_FLOAT, BT_SAMPLER, BT_STRUCT };
struct type_t {
  base_t b = BT_VOID;
  int n = 1;    // components of a scalar/vector
  int st = -1;  // struct index
};

struct semantic_t {
  bool has = false;
  std::string name;  // lower case
  uint32_t index = 0;
  int line = 0;
};

struct member_t {
  std::string name;
  type_t t;
  semantic_t sem;
};
struct struct_t {
  std::string name;
  std::vector<member_t> members;
};

// ---------------------------------------------------------------- expressions
enum node_kind { N_NUM, N_BOOL, N_ID, N_MEMBER, N_CALL, N_BIN, N_NEG };
struct node {
  node_kind k;
  std::string s;  // identifier, member, function name or operator
  double num = 0.0;
  bool is_float = false;
  int line = 0;
  std::vector<std::unique_ptr<node>> c;
};
using node_ptr = std::unique_ptr<node>;

// A variable: a register (scalar/vector), registers per member (struct), a constant slot
// (uniform) or a sampler slot.
enum var_kind { V_REG, V_STRUCT, V_UNIFORM, V_SAMPLER };
struct var_t {
  std::string name;
  var_kind kind;
  type_t t;
  int reg = -1;
  std::vector<int> mregs;
  int slot = -1;
};

constexpr uint8_t SWZ_ID = 0xE4;  // x y z w
inline uint8_t swz_comp(uint8_t s, int k) { return static_cast<uint8_t>((s >> (2 * k)) & 3); }
inline uint8_t swz_splat(int c) { return static_cast<uint8_t>(c | (c << 2) | (c << 4) | (c << 6)); }

// A value: an operand (register or constant) read through a swizzle.
struct val_t {
  type_t t;
  uint8_t opnd = 0;
  uint8_t swz = SWZ_ID;
  int inst = -1;  // the instruction that just produced this temporary (identity swizzle), else -1
  const var_t* var = nullptr;  // struct or sampler variable
};

// An assignment target: register `reg`, destination components comps[0..n).
struct lval_t {
  type_t t;
  int reg;
  int n;
  uint8_t comps[4];
};

class compiler {
public:
  explicit compiler(std::vector<token> t) : toks_(std::move(t)) {}

  void run(sasl_module& out) {
    std::memset(&out.prog, 0, sizeof out.prog);
    prog_ = &out.prog;
    mod_ = &out;
    scopes_.emplace_back();  // globals
    // functions at file scope: an identifier pair followed by '(' outside braces
    int nfunc = 0, brace = 0;
    for (size_t i = 0; i + 2 < toks_.size(); ++i) {
      const token& t = toks_[i];
      if (t.k == TK_PUNCT && t.s == "{") ++brace;
      if (t.k == TK_PUNCT && t.s == "}") --brace;
      if (brace == 0 && t.k == TK_ID && toks_[i + 1].k == TK_ID && toks_[i + 2].k == TK_PUNCT && toks_[i + 2].s == "(")
        ++nfunc;
    }
    if (nfunc > 1) raise(SASL_UNSUPPORTED, 1, "more than one function (helper functions are not supported)");
    bool have_entry = false;
    while (peek().k != TK_EOF) {
      if (is("struct")) {
        parse_struct();
        continue;
      }
      const int line = peek().line;
      type_t t = parse_type();
      const std::string name = expect_id();
      if (is("(")) {
        if (have_entry) raise(SASL_UNSUPPORTED, line, "more than one function (helper functions are not supported)");
        have_entry = true;
        parse_entry(t, name, line);
        continue;
      }
      parse_globals(t, name, line);
    }
    if (!have_entry) raise(SASL_ERROR, peek().line, "no entry function");
    out.prog.ninst = static_cast<uint32_t>(ninst_);
    out.prog.nregs = static_cast<uint32_t>(max_reg_);
  }

private:
  // ---- token helpers
  const token& peek(int o = 0) const { return toks_[std::min(pos_ + o, toks_.size() - 1)]; }
  bool is(const char* s, int o = 0) const {
    const token& t = peek(o);
    return (t.k == TK_PUNCT || t.k == TK_ID) && t.s == s;
  }
  bool accept(const char* s) {
    if (!is(s)) return false;
    ++pos_;
    return true;
  }
  void expect(const char* s) {
    if (!accept(s)) raise(SASL_ERROR, peek().line, std::string("expected '") + s + "' near '" + peek().s + "'");
  }
  std::string expect_id() {
    if (peek().k != TK_ID) raise(SASL_ERROR, peek().line, "expected an identifier near '" + peek().s + "'");
    return toks_[pos_++].s;
  }

  // ---- types
  bool builtin_type(const std::string& s, type_t& t, int line) const {
    static const char* const bases[] = {"float", "int", "bool", "half", "uint", "double"};
    if (s == "void") {
      t = type_t{BT_VOID, 0, -1};
      return true;
    }
    if (s == "sampler" || s == "sampler2D") {
      t = type_t{BT_SAMPLER, 1, -1};
      return true;
    }
    if (s == "sampler1D" || s == "sampler3D" || s == "samplerCUBE")
      raise(SASL_UNSUPPORTED, line, s + " (only 2D samplers are sampled by the pixel shaders)");
    for (const char* b : bases) {
      const size_t bl = std::strlen(b);
      if (s.compare(0, bl, b) != 0) continue;
      const std::string rest = s.substr(bl);
      int n = 1;
      if (rest.size() == 1 && rest[0] >= '1' && rest[0] <= '4') n = rest[0] - '0';
      else if (rest.size() == 3 && rest[1] == 'x') raise(SASL_UNSUPPORTED, line, "matrix types are not supported");
      else if (!rest.empty()) return false;
      const std::string base(b);
      if (base == "uint" || base == "double") raise(SASL_UNSUPPORTED, line, base + " is not supported");
      t.b = base == "int" ? BT_INT : base == "bool" ? BT_BOOL : BT_FLOAT;
      t.n = n;
      t.st = -1;
      return true;
    }
    return false;
  }
  bool is_type_start() const {
    if (peek().k != TK_ID) return false;
    type_t t;
    if (builtin_type(peek().s, t, peek().line)) return true;
    if (peek().s == "const" || peek().s == "uniform" || peek().s == "static") return true;
    return find_struct(peek().s) >= 0;
  }
  int find_struct(const std::string& n) const {
    for (size_t i = 0; i < structs_.size(); ++i)
      if (structs_[i].name == n) return static_cast<int>(i);
    return -1;
  }
  type_t parse_type() {
    while (is("const") || is("uniform") || is("in")) {
      if (is("in") && !(peek(1).k == TK_ID)) break;
      ++pos_;
    }
    if (is("static")) raise(SASL_UNSUPPORTED, peek().line, "static variables are not supported");
    if (is("out") || is("inout")) raise(SASL_UNSUPPORTED, peek().line, "out parameters are not supported");
    const token& tk = peek();
    if (tk.k != TK_ID) raise(SASL_ERROR, tk.line, "expected a type near '" + tk.s + "'");
    type_t t;
    if (builtin_type(tk.s, t, tk.line)) {
      ++pos_;
      return t;
    }
    const int si = find_struct(tk.s);
    if (si < 0) raise(SASL_ERROR, tk.line, "unknown type '" + tk.s + "'");
    ++pos_;
    return type_t{BT_STRUCT, 0, si};
  }
  // ": NAME", ": NAME(i)" or ": NAMEi" (semantic_value, shader/constants.h:52-83)
  semantic_t parse_semantic() {
    semantic_t s;
    if (!accept(":")) return s;
    s.has = true;
    s.line = peek().line;
    std::string id = expect_id();
    size_t d = id.size();
    while (d > 0 && std::isdigit(static_cast<unsigned char>(id[d - 1]))) --d;
    if (d < id.size()) {
      s.index = static_cast<uint32_t>(std::strtoul(id.c_str() + d, nullptr, 10));
      id = id.substr(0, d);
    }
    if (accept("(")) {
      if (peek().k != TK_NUM || peek().is_float) raise(SASL_ERROR, peek().line, "semantic index");
      s.index = static_cast<uint32_t>(peek().num);
      ++pos_;
      expect(")");
    }
    for (auto& ch : id) ch = static_cast<char>(std::tolower(static_cast<unsigned char>(ch)));
    s.name = id;
    return s;
  }

  void parse_struct() {
    const int line = peek().line;
    expect("struct");
    struct_t st;
    st.name = expect_id();
    if (find_struct(st.name) >= 0) raise(SASL_ERROR, line, "struct '" + st.name + "' redefined");
    expect("{");
    while (!accept("}")) {
      type_t t = parse_type();
      if (t.b == BT_STRUCT) raise(SASL_UNSUPPORTED, peek().line, "nested structs are not supported");
      if (t.b == BT_SAMPLER || t.b == BT_VOID) raise(SASL_ERROR, peek().line, "invalid member type");
      do {
        member_t m;
        m.name = expect_id();
        if (is("[")) raise(SASL_UNSUPPORTED, peek().line, "arrays are not supported");
        m.t = t;
        m.sem = parse_semantic();
        for (auto& o : st.members)
          if (o.name == m.name) raise(SASL_ERROR, line, "member '" + m.name + "' redefined");
        st.members.push_back(m);
      } while (accept(","));
      expect(";");
    }
    expect(";");
    structs_.push_back(st);
  }

  // global uniforms and samplers (reflection_impl::add_global_var, reflection_impl.cpp:120-140)
  void parse_globals(type_t t, std::string name, int line) {
    for (;;) {
      if (is("[")) raise(SASL_UNSUPPORTED, line, "arrays are not supported");
      if (is(":")) raise(SASL_UNSUPPORTED, line, "global variables with semantics are not supported");
      if (is("=")) raise(SASL_UNSUPPORTED, line, "initialised global variables are not supported");
      if (lookup(name)) raise(SASL_ERROR, line, "'" + name + "' redefined");
      var_t v;
      v.name = name;
      v.t = t;
      if (t.b == BT_SAMPLER) {
        if (mod_->samplers.size() >= SASL_MAX_SAMPLERS) raise(SASL_UNSUPPORTED, line, "more than 4 samplers");
        v.kind = V_SAMPLER;
        v.slot = static_cast<int>(mod_->samplers.size());
        mod_->samplers.push_back(name);
      } else if (t.b == BT_FLOAT || t.b == BT_INT) {
        v.kind = V_UNIFORM;
        v.slot = new_const();
        mod_->uniforms.push_back(sasl_uniform{name, static_cast<uint8_t>(v.slot), static_cast<uint8_t>(t.n),
                                              static_cast<uint8_t>(t.b == BT_INT), 0});
      } else {
        raise(SASL_UNSUPPORTED, line, "global variables must be float/int scalars or vectors, or samplers");
      }
      scopes_[0].push_back(v);
      if (!accept(",")) break;
      line = peek().line;
      name = expect_id();
    }
    expect(";");
  }

  // ---- registers and constants
  int new_reg(int line) {
    if (next_reg_ >= SASL_MAX_REGS) raise(SASL_UNSUPPORTED, line, "the shader needs more than 32 registers");
    const int r = next_reg_++;
    max_reg_ = std::max(max_reg_, next_reg_);
    return r;
  }
  int new_const() {
    if (nconst_ >= SASL_MAX_CONSTS) raise(SASL_UNSUPPORTED, peek().line, "more than 32 constants and uniforms");
    return nconst_++;
  }
  uint8_t literal(float v) {
    for (int i = 0; i < nconst_; ++i) {
      if (!is_literal_[i]) continue;
      float c = prog_->consts[i][0];
      if (std::memcmp(&c, &v, 4) == 0) return static_cast<uint8_t>(SASL_CONST + i);
    }
    const int s = new_const();
    is_literal_[s] = true;
    for (int k = 0; k < 4; ++k) prog_->consts[s][k] = v;
    return static_cast<uint8_t>(SASL_CONST + s);
  }

  int emit(uint8_t op, int dst, uint8_t wmask, uint8_t imm, uint8_t a = 0, uint8_t sa = SWZ_ID, uint8_t b = 0,
           uint8_t sb = SWZ_ID, uint8_t c = 0, uint8_t sc = SWZ_ID, int line = 0) {
    if (ninst_ >= SASL_MAX_INST) raise(SASL_UNSUPPORTED, line ? line : peek().line, "the shader is longer than 256 instructions");
    sasl_inst& in = prog_->code[ninst_];
    in.op = op;
    in.dst = static_cast<uint8_t>(dst);
    in.wmask = wmask;
    in.imm = imm;
    in.a = a;
    in.b = b;
    in.c = c;
    in.sa = sa;
    in.sb = sb;
    in.sc = sc;
    in.target = 0;
    return ninst_++;
  }

  // ---- scopes
  const var_t* lookup(const std::string& n) const {
    for (size_t s = scopes_.size(); s-- > 0;)
      for (const auto& v : scopes_[s])
        if (v.name == n) return &v;
    return nullptr;
  }
  void declare(var_t v, int line) {
    for (const auto& o : scopes_.back())
      if (o.name == v.name) raise(SASL_ERROR, line, "'" + v.name + "' redefined");
    scopes_.back().push_back(std::move(v));
  }
  struct scope_guard {
    compiler& c;
    int saved_reg;
    explicit scope_guard(compiler& cc) : c(cc), saved_reg(cc.next_reg_) { c.scopes_.emplace_back(); }
    ~scope_guard() {
      c.scopes_.pop_back();
      c.next_reg_ = saved_reg;
    }
  };

  // ---- entry function (reflector.cpp:158-205)
  struct io_entry {
    uint32_t index;
    int reg;
    int ncomp;
  };
  // The reference keeps semantics in an array ordered by lower_bound insertion
  // (reflection_impl.cpp:71-118); with TEXCOORD inputs and COLOR outputs only, that is index order.
  static void insert_sorted(std::vector<io_entry>& v, io_entry e, int line, const char* what) {
    size_t i = 0;
    while (i < v.size() && v[i].index < e.index) ++i;
    if (i < v.size() && v[i].index == e.index) raise(SASL_ERROR, line, std::string("duplicate ") + what + " semantic");
    v.insert(v.begin() + static_cast<long>(i), e);
  }
  void add_input(const semantic_t& s, const type_t& t, int reg, int line) {
    if (!s.has) raise(SASL_ERROR, line, "entry parameter without a semantic");
    if (s.name != "texcoord")
      raise(SASL_UNSUPPORTED, line, "pixel-shader inputs are TEXCOORD(n) only (psinput_semantic_usage)");
    if (t.b != BT_FLOAT) raise(SASL_UNSUPPORTED, line, "pixel-shader inputs must be float scalars or vectors");
    insert_sorted(inputs_, io_entry{s.index, reg, t.n}, line, "input");
  }
  void add_output(const semantic_t& s, const type_t& t, int line) {
    if (!s.has) raise(SASL_ERROR, line, "entry result without a semantic");
    if (s.name == "depth" || s.name == "sv_depth") raise(SASL_UNSUPPORTED, line, "DEPTH outputs are not supported");
    if (s.name != "color" && s.name != "sv_target")
      raise(SASL_UNSUPPORTED, line, "pixel-shader outputs are COLOR(n)/SV_Target(n) only (psoutput_semantic_usage)");
    if (t.b != BT_FLOAT) raise(SASL_UNSUPPORTED, line, "pixel-shader outputs must be float scalars or vectors");
    if (s.index >= 8) raise(SASL_ERROR, line, "render target index");
    insert_sorted(outputs_, io_entry{s.index, -1, t.n}, line, "output");
  }

  void parse_entry(type_t rt, const std::string& name, int line) {
    (void)name;
    expect("(");
    scopes_.emplace_back();  // parameters
    if (!is(")")) {
      do {
        const int pl = peek().line;
        type_t t = parse_type();
        var_t v;
        v.name = expect_id();
        v.t = t;
        semantic_t s = parse_semantic();
        if (t.b == BT_STRUCT) {
          if (s.has) raise(SASL_ERROR, pl, "struct parameter with a semantic");
          v.kind = V_STRUCT;
          for (const member_t& m : structs_[t.st].members) {
            const int r = new_reg(pl);
            v.mregs.push_back(r);
            add_input(m.sem, m.t, r, pl);
          }
        } else if (t.b == BT_FLOAT || t.b == BT_INT || t.b == BT_BOOL) {
          v.kind = V_REG;
          v.reg = new_reg(pl);
          add_input(s, t, v.reg, pl);
        } else {
          raise(SASL_ERROR, pl, "invalid parameter type");
        }
        declare(v, pl);
      } while (accept(","));
    }
    expect(")");
    semantic_t rs = parse_semantic();
    ret_t_ = rt;
    if (rt.b == BT_STRUCT) {
      if (rs.has) raise(SASL_ERROR, line, "struct result with a semantic");
      for (const member_t& m : structs_[rt.st].members) add_output(m.sem, m.t, line);
    } else if (rt.b == BT_FLOAT) {
      add_output(rs, rt, line);
    } else {
      raise(SASL_UNSUPPORTED, line, "the entry must return a float vector or a struct");
    }
    if (inputs_.size() > SASL_MAX_INPUTS) raise(SASL_UNSUPPORTED, line, "more than 8 inputs");
    if (outputs_.size() > SASL_MAX_OUTPUTS) raise(SASL_UNSUPPORTED, line, "more than 8 outputs");
    prog_->ninputs = static_cast<uint32_t>(inputs_.size());
    for (size_t i = 0; i < inputs_.size(); ++i)
      prog_->inputs[i] = sasl_input{static_cast<uint8_t>(inputs_[i].reg), static_cast<uint8_t>(i),
                                    static_cast<uint8_t>(inputs_[i].ncomp), 0};
    expect("{");
    returned_ = false;
    while (!accept("}")) {
      if (returned_) raise(SASL_UNSUPPORTED, peek().line, "statements after return");
      statement(true);
    }
    if (!returned_) raise(SASL_ERROR, peek().line, "the entry function does not return");
    scopes_.pop_back();
  }

  // ---- statements (single pass: code is emitted as statements are parsed)
  void statement(bool top) {
    const int line = peek().line;
    if (accept(";")) return;
    if (is("{")) {
      ++pos_;
      scope_guard g(*this);
      while (!accept("}")) statement(false);
      return;
    }
    if (is("if")) return if_statement();
    if (is("while")) return while_statement();
    if (is("do")) return do_statement();
    if (is("for")) return for_statement();
    if (accept("break")) {
      if (loop_depth_ == 0) raise(SASL_ERROR, line, "break outside a loop");
      emit(SOP_BREAK, 0, 0, 0);
      expect(";");
      return;
    }
    if (is("continue")) raise(SASL_UNSUPPORTED, line, "continue (unimplemented in the reference's SIMD code generator)");
    if (is("switch")) raise(SASL_UNSUPPORTED, line, "switch (unimplemented in the reference's SIMD code generator)");
    if (is("discard") || is("clip")) raise(SASL_UNSUPPORTED, line, "discard is not supported");
    if (accept("return")) {
      if (!top || depth_ != 0) raise(SASL_UNSUPPORTED, line, "return inside a branch or loop");
      return_statement(line);
      return;
    }
    if (is_type_start() && !(peek(1).k == TK_PUNCT && peek(1).s == "(")) {
      declaration();
      expect(";");
      return;
    }
    simple_statement();
    expect(";");
  }

  void declaration() {
    const int line = peek().line;
    type_t t = parse_type();
    if (t.b == BT_SAMPLER || t.b == BT_VOID) raise(SASL_ERROR, line, "invalid local variable type");
    do {
      const int dl = peek().line;
      var_t v;
      v.name = expect_id();
      v.t = t;
      if (is("[")) raise(SASL_UNSUPPORTED, dl, "arrays are not supported");
      if (t.b == BT_STRUCT) {
        v.kind = V_STRUCT;
        for (size_t m = 0; m < structs_[t.st].members.size(); ++m) v.mregs.push_back(new_reg(dl));
        if (is("=")) raise(SASL_UNSUPPORTED, dl, "struct initialisers are not supported");
        declare(v, dl);
        continue;
      }
      v.kind = V_REG;
      v.reg = new_reg(dl);
      if (accept("=")) {
        const int mark = next_reg_;
        node_ptr e = expression();
        val_t x = gen(e.get());
        lval_t L{t, v.reg, t.n, {0, 1, 2, 3}};
        store(L, x, dl);
        next_reg_ = mark;
      }
      declare(v, dl);  // visible after its initialiser
    } while (accept(","));
  }

  // assignment or a call statement
  void simple_statement() {
    const int line = peek().line;
    node_ptr lhs = expression();
    static const char* const ops[] = {"=", "+=", "-=", "*=", "/=", "%="};
    const char* op = nullptr;
    for (const char* o : ops)
      if (is(o)) op = o;
    if (!op) {
      if (is("++") || is("--")) raise(SASL_UNSUPPORTED, line, "++/-- (unary operators are unimplemented in the reference's SIMD code generator)");
      if (lhs->k != N_CALL) raise(SASL_ERROR, line, "expression statement has no effect");
      const int mark = next_reg_;
      gen(lhs.get());
      next_reg_ = mark;
      return;
    }
    ++pos_;
    node_ptr rhs = expression();
    emit_assign(lhs.get(), op, rhs.get(), line);
  }
  void emit_assign(const node* lhs, const std::string& op, const node* rhs, int line) {
    const int mark = next_reg_;
    lval_t L = genl(lhs);
    val_t x;
    if (op == "=") {
      x = gen(rhs);
    } else {
      val_t cur = gen(lhs);  // the target's current value
      val_t r = gen(rhs);
      x = arith(op.substr(0, 1), cur, r, line);
    }
    store(L, x, line);
    next_reg_ = mark;
  }

  void return_statement(int line) {
    if (ret_t_.b == BT_STRUCT) {
      node_ptr e = expression();
      expect(";");
      if (e->k != N_ID) raise(SASL_UNSUPPORTED, line, "return of a struct must name a variable");
      const var_t* v = lookup(e->s);
      if (!v || v->kind != V_STRUCT || v->t.st != ret_t_.st) raise(SASL_ERROR, line, "returned value has the wrong type");
      const struct_t& st = structs_[ret_t_.st];
      for (size_t m = 0; m < st.members.size(); ++m)
        for (auto& o : outputs_)
          if (o.index == st.members[m].sem.index) o.reg = v->mregs[m];
    } else {
      node_ptr e = expression();
      expect(";");
      val_t x = gen(e.get());
      if (x.t.b == BT_STRUCT || x.t.b == BT_SAMPLER) raise(SASL_ERROR, line, "returned value has the wrong type");
      // the returned value is read from a register holding it in components 0..n-1
      int reg;
      if (x.opnd < SASL_CONST && x.swz == SWZ_ID && x.t.n >= ret_t_.n) {
        reg = x.opnd;
      } else {
        reg = new_reg(line);
        lval_t L{ret_t_, reg, ret_t_.n, {0, 1, 2, 3}};
        store(L, x, line);
      }
      outputs_[0].reg = reg;
    }
    prog_->noutputs = static_cast<uint32_t>(outputs_.size());
    for (size_t i = 0; i < outputs_.size(); ++i)
      prog_->outputs[i] = sasl_output{static_cast<uint8_t>(outputs_[i].reg), static_cast<uint8_t>(outputs_[i].index),
                                      static_cast<uint8_t>(outputs_[i].ncomp), 0};
    returned_ = true;
  }

  void push_depth(int line) {
    if (++depth_ > SASL_MAX_DEPTH) raise(SASL_UNSUPPORTED, line, "if/loop nesting deeper than 8");
  }

  // if_statement (cg_simd.cpp:144-193)
  void if_statement() {
    const int line = peek().line;
    expect("if");
    expect("(");
    const uint8_t c = condition(line);
    expect(")");
    push_depth(line);
    const int i_if = emit(SOP_IF, 0, 0, 0, c, SWZ_ID);
    statement(false);
    int i_else = -1;
    if (accept("else")) {
      i_else = emit(SOP_ELSE, 0, 0, 0);
      statement(false);
    }
    const int i_end = emit(SOP_ENDIF, 0, 0, 0);
    prog_->code[i_if].target = static_cast<uint16_t>(i_else >= 0 ? i_else : i_end);
    if (i_else >= 0) prog_->code[i_else].target = static_cast<uint16_t>(i_end);
    --depth_;
  }
  // the condition's value in a register or constant, component x (temporaries released)
  uint8_t condition(int line) {
    const int mark = next_reg_;
    node_ptr e = expression();
    val_t x = gen(e.get());
    if (x.t.b == BT_STRUCT || x.t.b == BT_SAMPLER || x.t.n != 1) raise(SASL_ERROR, line, "the condition must be a scalar");
    uint8_t opnd = x.opnd;
    if (swz_comp(x.swz, 0) != 0) {
      const int t = new_reg(line);
      emit(SOP_MOV, t, 1, 0, x.opnd, x.swz);
      opnd = static_cast<uint8_t>(t);
    }
    next_reg_ = mark;  // read by the next instruction only
    return opnd;
  }

  // while_statement (cg_simd.cpp:195-229)
  void while_statement() {
    const int line = peek().line;
    expect("while");
    push_depth(line);
    emit(SOP_LOOP, 0, 0, 0);
    const int cond_start = ninst_;
    expect("(");
    const uint8_t c = condition(line);
    expect(")");
    const int i_test = emit(SOP_LTEST, 0, 0, 0, c, SWZ_ID);
    ++loop_depth_;
    statement(false);
    --loop_depth_;
    emit(SOP_LSAVE, 0, 0, 0);
    const int i_jmp = emit(SOP_JMP, 0, 0, 0);
    prog_->code[i_jmp].target = static_cast<uint16_t>(cond_start);
    const int i_end = emit(SOP_LEND, 0, 0, 0);
    prog_->code[i_test].target = static_cast<uint16_t>(i_end);
    --depth_;
  }
  // dowhile_statement (cg_simd.cpp:231-267)
  void do_statement() {
    const int line = peek().line;
    expect("do");
    push_depth(line);
    emit(SOP_LOOP, 0, 0, 0);
    const int body = emit(SOP_DBODY, 0, 0, 0);
    ++loop_depth_;
    statement(false);
    --loop_depth_;
    emit(SOP_LSAVE, 0, 0, 0);
    expect("while");
    expect("(");
    const uint8_t c = condition(line);
    expect(")");
    expect(";");
    const int i_test = emit(SOP_DTEST, 0, 0, 0, c, SWZ_ID);
    prog_->code[i_test].target = static_cast<uint16_t>(body);
    emit(SOP_LEND, 0, 0, 0);
    --depth_;
  }
  // for_statement (cg_simd.cpp:269-339): the iteration runs after the body, then the mask is saved
  void for_statement() {
    const int line = peek().line;
    expect("for");
    expect("(");
    push_depth(line);
    scope_guard g(*this);
    emit(SOP_LOOP, 0, 0, 0);
    if (!accept(";")) {
      if (is_type_start() && !(peek(1).k == TK_PUNCT && peek(1).s == "(")) declaration();
      else simple_statement();
      expect(";");
    }
    const int cond_start = ninst_;
    int i_test;
    if (is(";")) {
      i_test = emit(SOP_LTEST, 0, 0, 1);
    } else {
      const uint8_t c = condition(line);
      i_test = emit(SOP_LTEST, 0, 0, 0, c, SWZ_ID);
    }
    expect(";");
    // the iteration statement is parsed now and emitted after the body
    node_ptr it_l, it_r;
    std::string it_op;
    int it_line = peek().line;
    if (!is(")")) {
      it_l = expression();
      static const char* const ops[] = {"=", "+=", "-=", "*=", "/=", "%="};
      for (const char* o : ops)
        if (is(o)) it_op = o;
      if (it_op.empty()) {
        if (is("++") || is("--")) raise(SASL_UNSUPPORTED, it_line, "++/-- (unary operators are unimplemented in the reference's SIMD code generator)");
        raise(SASL_UNSUPPORTED, it_line, "the iteration must be an assignment");
      }
      ++pos_;
      it_r = expression();
    }
    expect(")");
    ++loop_depth_;
    statement(false);
    --loop_depth_;
    if (it_l) emit_assign(it_l.get(), it_op, it_r.get(), it_line);
    emit(SOP_LSAVE, 0, 0, 0);
    const int i_jmp = emit(SOP_JMP, 0, 0, 0);
    prog_->code[i_jmp].target = static_cast<uint16_t>(cond_start);
    const int i_end = emit(SOP_LEND, 0, 0, 0);
    prog_->code[i_test].target = static_cast<uint16_t>(i_end);
    --depth_;
  }

  // ---- expression parsing (C precedence)
  node_ptr mk(node_kind k, int line, std::string s = "") {
    node_ptr n(new node);
    n->k = k;
    n->line = line;
    n->s = std::move(s);
    return n;
  }
  node_ptr expression() {
    node_ptr e = binary(0);
    if (is("?")) raise(SASL_UNSUPPORTED, peek().line, "?: (unimplemented in the reference's SIMD code generator)");
    return e;
  }
  static int prec(const std::string& s) {
    if (s == "||") return 1;
    if (s == "&&") return 2;
    if (s == "|") return 3;
    if (s == "^") return 4;
    if (s == "&") return 5;
    if (s == "==" || s == "!=") return 6;
    if (s == "<" || s == ">" || s == "<=" || s == ">=") return 7;
    if (s == "<<" || s == ">>") return 8;
    if (s == "+" || s == "-") return 9;
    if (s == "*" || s == "/" || s == "%") return 10;
    return -1;
  }
  node_ptr binary(int min_prec) {
    node_ptr lhs = unary();
    for (;;) {
      const token& t = peek();
      if (t.k != TK_PUNCT) break;
      const int p = prec(t.s);
      if (p < 0 || p < min_prec) break;
      const std::string op = t.s;
      const int line = t.line;
      ++pos_;
      node_ptr rhs = binary(p + 1);
      node_ptr b = mk(N_BIN, line, op);
      b->c.push_back(std::move(lhs));
      b->c.push_back(std::move(rhs));
      lhs = std::move(b);
    }
    return lhs;
  }
  node_ptr unary() {
    const int line = peek().line;
    if (accept("-")) {
      node_ptr n = mk(N_NEG, line);
      n->c.push_back(unary());
      return n;
    }
    if (accept("+")) return unary();
    if (is("!") || is("~") || is("++") || is("--"))
      raise(SASL_UNSUPPORTED, line, "unary " + peek().s + " (unimplemented in the reference's SIMD code generator)");
    return postfix();
  }
  node_ptr postfix() {
    node_ptr e = primary();
    for (;;) {
      const int line = peek().line;
      if (accept(".")) {
        node_ptr m = mk(N_MEMBER, line, expect_id());
        m->c.push_back(std::move(e));
        e = std::move(m);
        continue;
      }
      if (is("[")) raise(SASL_UNSUPPORTED, line, "indexing is not supported");
      break;
    }
    return e;
  }
  node_ptr primary() {
    const token& t = peek();
    if (t.k == TK_NUM) {
      node_ptr n = mk(N_NUM, t.line);
      n->num = t.num;
      n->is_float = t.is_float;
      ++pos_;
      return n;
    }
    if (accept("(")) {
      if (is_type_start() && peek(1).k == TK_PUNCT && peek(1).s == ")")
        raise(SASL_UNSUPPORTED, t.line, "casts (unimplemented in the reference's SIMD code generator)");
      node_ptr e = expression();
      expect(")");
      return e;
    }
    if (t.k == TK_ID) {
      const int line = t.line;
      std::string id = t.s;
      ++pos_;
      if (id == "true" || id == "false") {
        node_ptr n = mk(N_BOOL, line);
        n->num = id == "true" ? 1.0 : 0.0;
        return n;
      }
      if (accept("(")) {
        node_ptr call = mk(N_CALL, line, id);
        if (!is(")")) {
          do call->c.push_back(expression());
          while (accept(","));
        }
        expect(")");
        return call;
      }
      return mk(N_ID, line, id);
    }
    raise(SASL_ERROR, t.line, "unexpected '" + t.s + "'");
  }

  // ---- code generation
  static bool numeric(const type_t& t) { return t.b == BT_FLOAT || t.b == BT_INT || t.b == BT_BOOL; }

  val_t temp_result(type_t t, int inst, int reg) {
    val_t v;
    v.t = t;
    v.opnd = static_cast<uint8_t>(reg);
    v.swz = SWZ_ID;
    v.inst = inst;
    return v;
  }
  // operand swizzle of x read as an n-component value (scalars broadcast)
  static uint8_t widen(const val_t& x, int n) {
    if (x.t.n == 1 && n > 1) return swz_splat(swz_comp(x.swz, 0));
    return x.swz;
  }
  static uint8_t mask_of(int n) { return static_cast<uint8_t>((1 << n) - 1); }

  val_t op1(uint8_t op, uint8_t imm, const val_t& a, type_t rt, int line) {
    const int t = new_reg(line);
    const int i = emit(op, t, mask_of(rt.n), imm, a.opnd, widen(a, rt.n), 0, SWZ_ID, 0, SWZ_ID, line);
    return temp_result(rt, i, t);
  }
  val_t op2(uint8_t op, uint8_t imm, const val_t& a, const val_t& b, type_t rt, int line) {
    const int t = new_reg(line);
    const int i = emit(op, t, mask_of(rt.n), imm, a.opnd, widen(a, rt.n), b.opnd, widen(b, rt.n), 0, SWZ_ID, line);
    return temp_result(rt, i, t);
  }
  int common_n(const val_t& a, const val_t& b, int line) const {
    if (a.t.n == b.t.n) return a.t.n;
    if (a.t.n == 1) return b.t.n;
    if (b.t.n == 1) return a.t.n;
    raise(SASL_ERROR, line, "vector sizes differ");
  }
  void need_numeric(const val_t& v, int line) const {
    if (!numeric(v.t)) raise(SASL_ERROR, line, "operand is not a number or vector");
  }

  val_t arith(const std::string& op, const val_t& a, const val_t& b, int line) {
    need_numeric(a, line);
    need_numeric(b, line);
    if (a.t.b == BT_BOOL || b.t.b == BT_BOOL) raise(SASL_UNSUPPORTED, line, "arithmetic on bools");
    type_t rt{(a.t.b == BT_FLOAT || b.t.b == BT_FLOAT) ? BT_FLOAT : BT_INT, common_n(a, b, line), -1};
    if (rt.b == BT_INT && (op == "/" || op == "%")) raise(SASL_UNSUPPORTED, line, "integer division");
    uint8_t o;
    if (op == "+") o = SOP_ADD;
    else if (op == "-") o = SOP_SUB;
    else if (op == "*") o = SOP_MUL;
    else if (op == "/") o = SOP_DIV;
    else if (op == "%") return op2(SOP_FN2, SFN_FMOD, a, b, rt, line);  // emit_mod: fmodf (cgs.cpp:718-749)
    else raise(SASL_UNSUPPORTED, line, "operator " + op);
    return op2(o, 0, a, b, rt, line);
  }

  val_t gen(const node* e) {
    const int line = e->line;
    switch (e->k) {
    case N_NUM: {
      val_t v;
      v.t = type_t{e->is_float ? BT_FLOAT : BT_INT, 1, -1};
      v.opnd = literal(static_cast<float>(e->num));
      v.swz = SWZ_ID;
      return v;
    }
    case N_BOOL: {
      val_t v;
      v.t = type_t{BT_BOOL, 1, -1};
      v.opnd = literal(static_cast<float>(e->num));
      return v;
    }
    case N_ID: {
      const var_t* var = lookup(e->s);
      if (!var) raise(SASL_ERROR, line, "unknown identifier '" + e->s + "'");
      val_t v;
      v.t = var->t;
      v.var = var;
      if (var->kind == V_REG) v.opnd = static_cast<uint8_t>(var->reg);
      else if (var->kind == V_UNIFORM) v.opnd = static_cast<uint8_t>(SASL_CONST + var->slot);
      return v;
    }
    case N_MEMBER: {
      val_t base = gen(e->c[0].get());
      if (base.t.b == BT_STRUCT) {
        const struct_t& st = structs_[base.t.st];
        for (size_t m = 0; m < st.members.size(); ++m)
          if (st.members[m].name == e->s) {
            val_t v;
            v.t = st.members[m].t;
            v.opnd = static_cast<uint8_t>(base.var->mregs[m]);
            return v;
          }
        raise(SASL_ERROR, line, "no member '" + e->s + "' in " + st.name);
      }
      need_numeric(base, line);
      uint8_t comps[4];
      const int n = parse_swizzle(e->s, base.t.n, comps, line);
      val_t v = base;
      v.inst = -1;
      v.t.n = n;
      uint8_t s = 0;
      for (int k = 0; k < 4; ++k) s |= static_cast<uint8_t>(swz_comp(base.swz, comps[k < n ? k : n - 1]) << (2 * k));
      v.swz = s;
      return v;
    }
    case N_NEG: {
      val_t a = gen(e->c[0].get());
      need_numeric(a, line);
      if (a.t.b == BT_BOOL) raise(SASL_UNSUPPORTED, line, "negation of a bool");
      return op1(SOP_NEG, 0, a, a.t, line);
    }
    case N_BIN: {
      const std::string& op = e->s;
      if (op == "&&" || op == "||")
        raise(SASL_UNSUPPORTED, line, op + " (unimplemented in the reference's SIMD code generator)");
      if (op == "&" || op == "|" || op == "^" || op == "<<" || op == ">>")
        raise(SASL_UNSUPPORTED, line, "bitwise operators are not supported");
      val_t a = gen(e->c[0].get());
      val_t b = gen(e->c[1].get());
      if (prec(op) == 6 || prec(op) == 7) {
        need_numeric(a, line);
        need_numeric(b, line);
        type_t rt{BT_BOOL, common_n(a, b, line), -1};
        uint8_t o = op == "<" ? SOP_CLT : op == ">" ? SOP_CGT : op == "<=" ? SOP_CLE : op == ">=" ? SOP_CGE : op == "==" ? SOP_CEQ : SOP_CNE;
        return op2(o, 0, a, b, rt, line);
      }
      return arith(op, a, b, line);
    }
    case N_CALL: return call(e);
    }
    raise(SASL_ERROR, line, "expression");
  }

  static int parse_swizzle(const std::string& s, int n, uint8_t* comps, int line) {
    if (s.empty() || s.size() > 4) raise(SASL_ERROR, line, "invalid swizzle '" + s + "'");
    const bool xyzw = std::strchr("xyzw", s[0]) != nullptr;
    for (size_t k = 0; k < s.size(); ++k) {
      const char* set = xyzw ? "xyzw" : "rgba";
      const char* f = std::strchr(set, s[k]);
      if (!f || !s[k]) raise(SASL_ERROR, line, "invalid swizzle '" + s + "'");
      const int c = static_cast<int>(f - set);
      if (c >= n) raise(SASL_ERROR, line, "swizzle '" + s + "' reads past the vector");
      comps[k] = static_cast<uint8_t>(c);
    }
    return static_cast<int>(s.size());
  }

  lval_t genl(const node* e) {
    const int line = e->line;
    if (e->k == N_ID) {
      const var_t* v = lookup(e->s);
      if (!v) raise(SASL_ERROR, line, "unknown identifier '" + e->s + "'");
      if (v->kind == V_UNIFORM || v->kind == V_SAMPLER) raise(SASL_ERROR, line, "'" + e->s + "' cannot be assigned");
      if (v->kind == V_STRUCT) raise(SASL_UNSUPPORTED, line, "whole-struct assignment is not supported");
      return lval_t{v->t, v->reg, v->t.n, {0, 1, 2, 3}};
    }
    if (e->k == N_MEMBER) {
      const node* o = e->c[0].get();
      if (o->k == N_ID) {
        const var_t* v = lookup(o->s);
        if (v && v->kind == V_STRUCT) {
          const struct_t& st = structs_[v->t.st];
          for (size_t m = 0; m < st.members.size(); ++m)
            if (st.members[m].name == e->s) return lval_t{st.members[m].t, v->mregs[m], st.members[m].t.n, {0, 1, 2, 3}};
          raise(SASL_ERROR, line, "no member '" + e->s + "' in " + st.name);
        }
      }
      lval_t base = genl(o);
      uint8_t comps[4];
      const int n = parse_swizzle(e->s, base.n, comps, line);
      lval_t L = base;
      L.n = n;
      L.t.n = n;
      for (int k = 0; k < n; ++k) {
        L.comps[k] = base.comps[comps[k]];
        for (int j = 0; j < k; ++j)
          if (L.comps[j] == L.comps[k]) raise(SASL_ERROR, line, "write mask names a component twice");
      }
      return L;
    }
    raise(SASL_ERROR, line, "the left side is not assignable");
  }

  // cgs_simd::store: a masked write of x into L (scalars broadcast, longer vectors truncated)
  void store(const lval_t& L, const val_t& x, int line) {
    need_numeric(x, line);
    if (x.t.n != 1 && x.t.n < L.n) raise(SASL_ERROR, line, "vector sizes differ");
    if (L.t.b == BT_INT && x.t.b == BT_FLOAT)
      raise(SASL_UNSUPPORTED, line, "float to int conversion (casts are unimplemented in the reference's SIMD code generator)");
    uint8_t wm = 0;
    for (int k = 0; k < L.n; ++k) wm |= static_cast<uint8_t>(1u << L.comps[k]);
    wm |= SASL_WM_MASKED;
    // component k of the value goes to destination component L.comps[k]
    auto remap = [&](uint8_t s) {
      uint8_t r = 0;
      for (int k = 0; k < L.n; ++k) {
        const int src = x.t.n == 1 ? 0 : k;
        r |= static_cast<uint8_t>(swz_comp(s, src) << (2 * L.comps[k]));
      }
      return r;
    };
    // a temporary computed per component by the instruction just emitted is written in place
    if (x.inst >= 0 && x.inst == ninst_ - 1 && x.opnd < SASL_CONST) {
      sasl_inst& in = prog_->code[x.inst];
      const bool per_component = in.op != SOP_CROSS && in.op != SOP_TEX && in.op != SOP_DOT;
      if (per_component && in.dst == x.opnd && (x.t.n == L.n || (x.t.n == 1 && in.op != SOP_DOT))) {
        in.dst = static_cast<uint8_t>(L.reg);
        in.wmask = wm;
        in.sa = remap(in.sa);
        in.sb = remap(in.sb);
        in.sc = remap(in.sc);
        return;
      }
      if (in.op == SOP_DOT && in.dst == x.opnd) {  // every component holds the sum
        in.dst = static_cast<uint8_t>(L.reg);
        in.wmask = wm;
        return;
      }
    }
    emit(SOP_MOV, L.reg, wm, 0, x.opnd, remap(x.swz), 0, SWZ_ID, 0, SWZ_ID, line);
  }

  // ---- calls: constructors and intrinsics (cg_impl.cpp:790-1240)
  val_t arg(const node* e, size_t i, int line) {
    if (i >= e->c.size()) raise(SASL_ERROR, line, e->s + ": too few arguments");
    return gen(e->c[i].get());
  }
  void nargs(const node* e, size_t n) const {
    if (e->c.size() != n) raise(SASL_ERROR, e->line, e->s + " takes " + std::to_string(n) + " arguments");
  }
  val_t fvec(const val_t& v, int line) const {
    if (!numeric(v.t)) raise(SASL_ERROR, line, "argument is not a number or vector");
    if (v.t.b == BT_BOOL) raise(SASL_UNSUPPORTED, line, "bool argument");
    return v;
  }
  static type_t ftype(int n) { return type_t{BT_FLOAT, n, -1}; }

  val_t call(const node* e) {
    const std::string& f = e->s;
    const int line = e->line;
    type_t ct;
    if (builtin_type(f, ct, line) && (ct.b == BT_FLOAT || ct.b == BT_INT || ct.b == BT_BOOL)) {
      // vector constructor: arguments fill the components in order (cg_impl.cpp:956-991)
      const int t = new_reg(line);
      int k = 0;
      int last = -1;
      for (size_t i = 0; i < e->c.size(); ++i) {
        val_t a = gen(e->c[i].get());
        need_numeric(a, line);
        if (ct.b == BT_INT && a.t.b == BT_FLOAT) raise(SASL_UNSUPPORTED, line, "float to int conversion");
        if (k + a.t.n > ct.n) raise(SASL_ERROR, line, f + ": too many components");
        uint8_t wm = 0, s = 0;
        for (int j = 0; j < a.t.n; ++j) {
          wm |= static_cast<uint8_t>(1u << (k + j));
          s |= static_cast<uint8_t>(swz_comp(a.swz, j) << (2 * (k + j)));
        }
        last = emit(SOP_MOV, t, wm, 0, a.opnd, s, 0, SWZ_ID, 0, SWZ_ID, line);
        k += a.t.n;
      }
      if (k != ct.n) raise(SASL_ERROR, line, f + ": " + std::to_string(k) + " components for " + std::to_string(ct.n));
      val_t v = temp_result(ct, -1, t);
      (void)last;
      return v;
    }
    struct fn1 {
      const char* name;
      uint8_t id;
    };
    static const fn1 unary_fns[] = {{"exp", SFN_EXP},   {"exp2", SFN_EXP2},   {"log", SFN_LOG},     {"log2", SFN_LOG2},
                                    {"log10", SFN_LOG10}, {"sin", SFN_SIN},   {"cos", SFN_COS},     {"tan", SFN_TAN},
                                    {"asin", SFN_ASIN}, {"acos", SFN_ACOS},   {"atan", SFN_ATAN},   {"sinh", SFN_SINH},
                                    {"cosh", SFN_COSH}, {"tanh", SFN_TANH},   {"ceil", SFN_CEIL},   {"floor", SFN_FLOOR},
                                    {"round", SFN_ROUND}, {"trunc", SFN_TRUNC}, {"rsqrt", SFN_RSQRT}};
    for (const fn1& u : unary_fns)
      if (f == u.name) {
        nargs(e, 1);
        val_t a = fvec(arg(e, 0, line), line);
        return op1(SOP_FN1, u.id, a, ftype(a.t.n), line);
      }
    if (f == "pow" || f == "fmod" || f == "ldexp") {
      nargs(e, 2);
      val_t a = fvec(arg(e, 0, line), line), b = fvec(arg(e, 1, line), line);
      const uint8_t id = f == "pow" ? SFN_POW : f == "fmod" ? SFN_FMOD : SFN_LDEXP;
      return op2(SOP_FN2, id, a, b, ftype(common_n(a, b, line)), line);
    }
    if (f == "abs" || f == "sqrt") {
      nargs(e, 1);
      val_t a = fvec(arg(e, 0, line), line);
      if (f == "sqrt" && a.t.b != BT_FLOAT) raise(SASL_UNSUPPORTED, line, "sqrt of an int");
      return op1(f == "abs" ? SOP_ABS : SOP_SQRT, 0, a, a.t, line);
    }
    if (f == "min" || f == "max") {  // select(v0 < v1, v0, v1) / select(v0 > v1, v0, v1)
      nargs(e, 2);
      val_t a = fvec(arg(e, 0, line), line), b = fvec(arg(e, 1, line), line);
      type_t rt{(a.t.b == BT_FLOAT || b.t.b == BT_FLOAT) ? BT_FLOAT : BT_INT, common_n(a, b, line), -1};
      return op2(f == "min" ? SOP_SLT : SOP_SGT, 0, a, b, rt, line);
    }
    if (f == "clamp" || f == "saturate") {  // emit_clamp (cgs.cpp:1860-1868)
      val_t v, lo, hi;
      if (f == "clamp") {
        nargs(e, 3);
        v = fvec(arg(e, 0, line), line);
        lo = fvec(arg(e, 1, line), line);
        hi = fvec(arg(e, 2, line), line);
      } else {
        nargs(e, 1);
        v = fvec(arg(e, 0, line), line);
        lo.t = hi.t = type_t{v.t.b, 1, -1};
        lo.opnd = literal(0.0f);
        hi.opnd = literal(1.0f);
      }
      type_t rt{v.t.b == BT_INT && lo.t.b == BT_INT && hi.t.b == BT_INT ? BT_INT : BT_FLOAT, v.t.n, -1};
      val_t r = op2(SOP_SGE, 0, v, lo, rt, line);
      return op2(SOP_SLE, 0, r, hi, rt, line);
    }
    if (f == "dot") {
      nargs(e, 2);
      val_t a = fvec(arg(e, 0, line), line), b = fvec(arg(e, 1, line), line);
      if (a.t.n != b.t.n) raise(SASL_ERROR, line, "dot: vector sizes differ");
      return dot(a, b, line);
    }
    if (f == "cross") {
      nargs(e, 2);
      val_t a = fvec(arg(e, 0, line), line), b = fvec(arg(e, 1, line), line);
      if (a.t.n != 3 || b.t.n != 3) raise(SASL_ERROR, line, "cross takes float3 arguments");
      return op2(SOP_CROSS, 0, a, b, ftype(3), line);
    }
    if (f == "length") {  // sqrt(dot(v, v))
      nargs(e, 1);
      val_t a = fvec(arg(e, 0, line), line);
      return op1(SOP_SQRT, 0, dot(a, a, line), ftype(1), line);
    }
    if (f == "distance") {  // sqrt(dot(d - s, d - s))
      nargs(e, 2);
      val_t s = fvec(arg(e, 0, line), line), d = fvec(arg(e, 1, line), line);
      val_t diff = op2(SOP_SUB, 0, d, s, ftype(common_n(s, d, line)), line);
      return op1(SOP_SQRT, 0, dot(diff, diff, line), ftype(1), line);
    }
    if (f == "normalize") {  // v / sqrt(dot(v, v))
      nargs(e, 1);
      val_t v = fvec(arg(e, 0, line), line);
      val_t len = op1(SOP_SQRT, 0, dot(v, v, line), ftype(1), line);
      return op2(SOP_DIV, 0, v, len, ftype(v.t.n), line);
    }
    if (f == "reflect") {  // i - (2 * dot(i, n)) * n
      nargs(e, 2);
      val_t i = fvec(arg(e, 0, line), line), n = fvec(arg(e, 1, line), line);
      val_t two;
      two.t = ftype(1);
      two.opnd = literal(2.0f);
      val_t dd = op2(SOP_MUL, 0, two, dot(i, n, line), ftype(1), line);
      val_t s = op2(SOP_MUL, 0, dd, n, ftype(n.t.n), line);
      return op2(SOP_SUB, 0, i, s, ftype(common_n(i, s, line)), line);
    }
    if (f == "lerp") {  // s + (d - s) * t
      nargs(e, 3);
      val_t s = fvec(arg(e, 0, line), line), d = fvec(arg(e, 1, line), line), t = fvec(arg(e, 2, line), line);
      const int n = common_n(s, d, line);
      val_t diff = op2(SOP_SUB, 0, d, s, ftype(n), line);
      val_t td = op2(SOP_MUL, 0, diff, t, ftype(common_n(diff, t, line)), line);
      return op2(SOP_ADD, 0, s, td, ftype(common_n(s, td, line)), line);
    }
    if (f == "frac") {  // abs(v) - floor(abs(v))
      nargs(e, 1);
      val_t v = fvec(arg(e, 0, line), line);
      val_t a = op1(SOP_ABS, 0, v, ftype(v.t.n), line);
      val_t fl = op1(SOP_FN1, SFN_FLOOR, a, ftype(v.t.n), line);
      return op2(SOP_SUB, 0, a, fl, ftype(v.t.n), line);
    }
    if (f == "rcp") {  // 1 / v
      nargs(e, 1);
      val_t v = fvec(arg(e, 0, line), line);
      val_t one;
      one.t = ftype(1);
      one.opnd = literal(1.0f);
      return op2(SOP_DIV, 0, one, v, ftype(v.t.n), line);
    }
    if (f == "ddx" || f == "ddy") {
      nargs(e, 1);
      val_t v = fvec(arg(e, 0, line), line);
      return op1(f == "ddx" ? SOP_DDX : SOP_DDY, 0, v, ftype(v.t.n), line);
    }
    if (f == "tex2D" || f == "tex2Dgrad") {
      nargs(e, f == "tex2D" ? 2 : 4);
      val_t s = gen(e->c[0].get());
      if (s.t.b != BT_SAMPLER || !s.var) raise(SASL_ERROR, line, f + ": the first argument must be a sampler");
      val_t c = fvec(arg(e, 1, line), line);
      if (c.t.n < 2) raise(SASL_ERROR, line, f + ": the coordinate needs two components");
      val_t dx, dy;
      if (f == "tex2D") {
        dx = op1(SOP_DDX, 0, c, ftype(c.t.n), line);
        dy = op1(SOP_DDY, 0, c, ftype(c.t.n), line);
      } else {
        dx = fvec(arg(e, 2, line), line);
        dy = fvec(arg(e, 3, line), line);
        if (dx.t.n < 2 || dy.t.n < 2) raise(SASL_ERROR, line, "tex2Dgrad: derivatives need two components");
      }
      const int t = new_reg(line);
      const int i = emit(SOP_TEX, t, mask_of(4), static_cast<uint8_t>(s.var->slot), c.opnd, c.swz, dx.opnd, dx.swz, dy.opnd,
                         dy.swz, line);
      return temp_result(ftype(4), i, t);
    }
    if (f == "tex2Dlod" || f == "tex2Dbias" || f == "tex2Dproj" || f.compare(0, 7, "texCUBE") == 0)
      raise(SASL_UNSUPPORTED, line, f + " in a pixel shader (unimplemented in the reference, sampler_api.cpp:20-50)");
    if (f == "mul") raise(SASL_UNSUPPORTED, line, "mul (matrices are not supported)");
    raise(SASL_UNSUPPORTED, line, "function '" + f + "' is not supported");
  }

  val_t dot(const val_t& a, const val_t& b, int line) {
    const int t = new_reg(line);
    const int i = emit(SOP_DOT, t, 1, static_cast<uint8_t>(a.t.n), a.opnd, a.swz, b.opnd, b.swz, 0, SWZ_ID, line);
    return temp_result(ftype(1), i, t);
  }

  std::vector<token> toks_;
  size_t pos_ = 0;
  sasl_program* prog_ = nullptr;
  sasl_module* mod_ = nullptr;
  std::vector<struct_t> structs_;
  std::vector<std::vector<var_t>> scopes_;
  std::vector<io_entry> inputs_, outputs_;
  type_t ret_t_;
  bool returned_ = false;
  int ninst_ = 0, next_reg_ = 0, max_reg_ = 0, nconst_ = 0;
  bool is_literal_[SASL_MAX_CONSTS] = {};
  int depth_ = 0, loop_depth_ = 0;
};

}  // namespace

int sasl_compile_ps(const char* src, size_t len, sasl_module& out, std::string& msg) {
  out = sasl_module{};
  try {
    compiler c(lex(src, len));
    c.run(out);
  } catch (const compile_error& e) {
    msg = e.msg;
    out = sasl_module{};
    return e.status;
  }
  msg.clear();
  return SASL_OK;
}

bool sasl_set_uniform(const sasl_module& m, sasl_program& prog, const std::string& name, const void* data, size_t size) {
  for (const sasl_uniform& u : m.uniforms) {
    if (u.name != name) continue;
    const size_t n = std::min<size_t>(u.ncomp, size / 4);
    for (size_t k = 0; k < n; ++k) {
      if (u.is_int) {
        int32_t v;
        std::memcpy(&v, static_cast<const uint8_t*>(data) + 4 * k, 4);
        prog.consts[u.cslot][k] = static_cast<float>(v);
      } else {
        std::memcpy(&prog.consts[u.cslot][k], static_cast<const uint8_t*>(data) + 4 * k, 4);
      }
    }
    return true;
  }
  return false;
}

}  // namespace srx
