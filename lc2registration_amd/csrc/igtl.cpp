// OpenIGTLink TRANSFORM streaming (include/lc2igtl.h): message packing written from the
// protocol's description, and a one-client TCP server on POSIX sockets.
#include "lc2igtl.h"

#include <arpa/inet.h>
#include <netinet/in.h>
#include <netinet/tcp.h>
#include <poll.h>
#include <sys/socket.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

namespace {

thread_local std::string g_err;

int fail(const std::string& m) {
    g_err = m;
    return -1;
}

// CRC-64 with the ECMA-182 polynomial, MSB first, initial value 0, no final xor.
uint64_t crc64(const unsigned char* p, size_t n) {
    static uint64_t table[256];
    static std::once_flag once;
    std::call_once(once, [] {
        for (int i = 0; i < 256; ++i) {
            uint64_t c = (uint64_t)i << 56;
            for (int k = 0; k < 8; ++k) c = (c & (1ULL << 63)) ? (c << 1) ^ 0x42F0E1EBA9EA3693ULL : c << 1;
            table[i] = c;
        }
    });
    uint64_t crc = 0;
    for (size_t i = 0; i < n; ++i) crc = table[((crc >> 56) ^ p[i]) & 0xFF] ^ (crc << 8);
    return crc;
}

void put16(unsigned char* p, uint16_t v) {
    p[0] = (unsigned char)(v >> 8);
    p[1] = (unsigned char)v;
}
void put32(unsigned char* p, uint32_t v) {
    for (int i = 0; i < 4; ++i) p[i] = (unsigned char)(v >> (24 - 8 * i));
}
void put64(unsigned char* p, uint64_t v) {
    for (int i = 0; i < 8; ++i) p[i] = (unsigned char)(v >> (56 - 8 * i));
}
void putf(unsigned char* p, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put32(p, u);
}

constexpr size_t kHeader = 58, kContent = 48, kExtHeader = 12, kMetaHeader = 2;

}  // namespace

struct lc2igtl_server {
    int listen_fd = -1;
    int client_fd = -1;
    int port = 0;
    unsigned next_id = 0;
    long long sent = 0;
    std::mutex mu;
};

extern "C" {

const char* lc2igtl_last_error(void) { return g_err.c_str(); }

int lc2igtl_pack_transform(const char* device_name, const float m16[16], double timestamp_s,
                           int header_version, unsigned message_id, unsigned char* out, size_t cap,
                           size_t* len) {
    if (!m16 || !len || (header_version != 1 && header_version != 2))
        return fail("lc2igtl_pack_transform: invalid argument");
    const size_t body = header_version == 2 ? kExtHeader + kContent + kMetaHeader : kContent;
    *len = kHeader + body;
    if (!out) return 0;
    if (cap < *len) return fail("lc2igtl_pack_transform: buffer too small");
    std::memset(out, 0, *len);
    unsigned char* h = out;
    put16(h, (uint16_t)header_version);
    std::memcpy(h + 2, "TRANSFORM", 9);  // type, char[12], zero padded
    if (device_name) std::strncpy(reinterpret_cast<char*>(h + 14), device_name, 20);
    const double sec = timestamp_s > 0.0 ? timestamp_s : 0.0;
    const uint64_t whole = (uint64_t)sec;
    const uint64_t frac = (uint64_t)((sec - (double)whole) * 4294967296.0);
    put64(h + 34, (whole << 32) | (frac & 0xFFFFFFFFULL));
    put64(h + 42, body);
    unsigned char* b = out + kHeader;
    unsigned char* c = b;
    if (header_version == 2) {
        put16(b, (uint16_t)kExtHeader);
        put16(b + 2, (uint16_t)kMetaHeader);
        put32(b + 4, 0);  // metadata bytes
        put32(b + 8, message_id);
        c = b + kExtHeader;
    }
    // rotation column by column, then translation
    const int order[12] = {0, 4, 8, 1, 5, 9, 2, 6, 10, 3, 7, 11};
    for (int i = 0; i < 12; ++i) putf(c + 4 * i, m16[order[i]]);
    if (header_version == 2) put16(c + kContent, 0);  // metadata index count
    put64(h + 50, crc64(b, body));
    return 0;
}

int lc2igtl_server_create(int port, int loopback_only, lc2igtl_server** out) {
    if (!out || port < 0 || port > 65535) return fail("lc2igtl_server_create: invalid argument");
    const int fd = ::socket(AF_INET, SOCK_STREAM, 0);
    if (fd < 0) return fail(std::string("socket: ") + std::strerror(errno));
    const int one = 1;
    (void)::setsockopt(fd, SOL_SOCKET, SO_REUSEADDR, &one, sizeof(one));
    sockaddr_in a{};
    a.sin_family = AF_INET;
    a.sin_port = htons((uint16_t)port);
    a.sin_addr.s_addr = htonl(loopback_only ? INADDR_LOOPBACK : INADDR_ANY);
    if (::bind(fd, reinterpret_cast<sockaddr*>(&a), sizeof(a)) < 0 || ::listen(fd, 1) < 0) {
        const std::string e = std::strerror(errno);
        ::close(fd);
        return fail("cannot create a server socket on port " + std::to_string(port) + ": " + e);
    }
    socklen_t al = sizeof(a);
    (void)::getsockname(fd, reinterpret_cast<sockaddr*>(&a), &al);
    auto* s = new lc2igtl_server();
    s->listen_fd = fd;
    s->port = ntohs(a.sin_port);
    *out = s;
    return 0;
}

int lc2igtl_port(lc2igtl_server* s) { return s ? s->port : -1; }

int lc2igtl_wait_client(lc2igtl_server* s, double timeout_ms) {
    if (!s) return fail("lc2igtl_wait_client: invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->client_fd >= 0) return 1;
    const auto t0 = std::chrono::steady_clock::now();
    while (true) {
        // poll in 200 ms steps, as the reference's WaitForConnection(200) loop
        pollfd p{s->listen_fd, POLLIN, 0};
        int step = 200;
        if (timeout_ms >= 0.0) {
            const double left = timeout_ms - std::chrono::duration<double, std::milli>(
                                                 std::chrono::steady_clock::now() - t0).count();
            if (left <= 0.0) return 0;
            step = left < 200.0 ? (int)left + 1 : 200;
        }
        const int r = ::poll(&p, 1, step);
        if (r < 0 && errno != EINTR) return fail(std::string("poll: ") + std::strerror(errno));
        if (r > 0) {
            const int c = ::accept(s->listen_fd, nullptr, nullptr);
            if (c >= 0) {
                const int one = 1;
                (void)::setsockopt(c, IPPROTO_TCP, TCP_NODELAY, &one, sizeof(one));
                s->client_fd = c;
                return 1;
            }
        }
    }
}

int lc2igtl_send_transform(lc2igtl_server* s, const char* device_name, const float m16[16]) {
    if (!s || !m16) return fail("lc2igtl_send_transform: invalid argument");
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->client_fd < 0) return fail("lc2igtl_send_transform: no client connected");
    unsigned char buf[kHeader + kExtHeader + kContent + kMetaHeader];
    size_t len = 0;
    const double now = std::chrono::duration<double>(
                           std::chrono::system_clock::now().time_since_epoch()).count();
    if (lc2igtl_pack_transform(device_name, m16, now, 2, s->next_id, buf, sizeof(buf), &len) < 0)
        return -1;
    size_t off = 0;
    while (off < len) {
        const ssize_t w = ::send(s->client_fd, buf + off, len - off, MSG_NOSIGNAL);
        if (w < 0) {
            if (errno == EINTR) continue;
            return fail(std::string("send: ") + std::strerror(errno));
        }
        off += (size_t)w;
    }
    ++s->next_id;
    ++s->sent;
    return 0;
}

long long lc2igtl_sent(lc2igtl_server* s) { return s ? s->sent : -1; }

void lc2igtl_server_destroy(lc2igtl_server* s) {
    if (!s) return;
    if (s->client_fd >= 0) ::close(s->client_fd);
    if (s->listen_fd >= 0) ::close(s->listen_fd);
    delete s;
}

}  // extern "C"
