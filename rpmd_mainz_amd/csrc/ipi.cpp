// ipi.cpp -- the i-PI socket client (src/ipi_driver.f90:1-254 over src/sockets.c) on the C-ABI:
// one GPU context serves positions from an i-PI server with the forces of full_forces.  Host code
// only (POSIX sockets).  The wire format is the i-PI protocol's: 12-byte space-padded headers,
// native-endian 4-byte integers and 8-byte doubles, the 3x3 cell as nine doubles.
#include <netdb.h>
#include <sys/socket.h>
#include <sys/un.h>
#include <unistd.h>

#include <cerrno>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/rpmd_mi355x.h"

namespace {

constexpr int MSGLEN = 12;

struct Fd {
    int fd = -1;
    ~Fd() {
        if (fd >= 0) close(fd);
    }
};

// the whole buffer, or false when the stream ends or breaks first
bool read_all(int fd, void* buf, size_t len) {
    char* p = static_cast<char*>(buf);
    while (len > 0) {
        const ssize_t n = recv(fd, p, len, 0);
        if (n < 0 && errno == EINTR) continue;
        if (n <= 0) return false;
        p += n;
        len -= (size_t)n;
    }
    return true;
}
bool write_all(int fd, const void* buf, size_t len) {
    const char* p = static_cast<const char*>(buf);
    while (len > 0) {
        // no SIGPIPE when the server has gone: the call fails instead
        const ssize_t n = send(fd, p, len, MSG_NOSIGNAL);
        if (n < 0 && errno == EINTR) continue;
        if (n <= 0) return false;
        p += n;
        len -= (size_t)n;
    }
    return true;
}
bool write_header(int fd, const char* h) {
    char m[MSGLEN];
    memset(m, ' ', MSGLEN);
    memcpy(m, h, strnlen(h, MSGLEN));
    return write_all(fd, m, MSGLEN);
}

// "prefix:host:port" as the reference splits it (src/ipi_driver.f90:90-97): the port after the last
// colon, the host between the first and the last; prefix "unix": the UNIX socket /tmp/ipi_<host>
// (src/sockets.c open_socket_), otherwise an IPv4 internet socket to host:port.  The reference holds
// the host in 12 characters; names of 12 or more are not cut here.
int connect_to(const char* address, int* out) {
    const std::string a(address);
    const size_t first = a.find(':'), last = a.rfind(':');
    if (first == std::string::npos || last == first) return RPMD_ERR_ARG;
    const std::string prefix = a.substr(0, first), host = a.substr(first + 1, last - first - 1),
                      port = a.substr(last + 1);
    int fd = -1;
    if (prefix == "unix") {
        sockaddr_un sa{};
        sa.sun_family = AF_UNIX;
        const std::string path = "/tmp/ipi_" + host;
        if (path.size() >= sizeof(sa.sun_path)) return RPMD_ERR_ARG;
        memcpy(sa.sun_path, path.c_str(), path.size() + 1);
        fd = socket(AF_UNIX, SOCK_STREAM, 0);
        if (fd < 0) return RPMD_ERR_SOCKET;
        if (connect(fd, reinterpret_cast<sockaddr*>(&sa), sizeof sa) < 0) {
            close(fd);
            return RPMD_ERR_SOCKET;
        }
    } else {
        addrinfo hints{}, *res = nullptr;
        hints.ai_family = AF_INET;
        hints.ai_socktype = SOCK_STREAM;
        if (getaddrinfo(host.c_str(), port.c_str(), &hints, &res) != 0 || !res) return RPMD_ERR_SOCKET;
        for (addrinfo* p = res; p; p = p->ai_next) {
            fd = socket(p->ai_family, p->ai_socktype, p->ai_protocol);
            if (fd < 0) continue;
            if (connect(fd, p->ai_addr, p->ai_addrlen) == 0) break;
            close(fd);
            fd = -1;
        }
        freeaddrinfo(res);
        if (fd < 0) return RPMD_ERR_SOCKET;
    }
    *out = fd;
    return 0;
}

}  // namespace

extern "C" int rpmd_ipi_run(rpmd_ctx* ctx, const char* address, const double* z, int na, int* served) {
    if (!ctx || !address || !z || na <= 0) return RPMD_ERR_ARG;
    if (served) *served = 0;
    Fd sock;
    if (int rc = connect_to(address, &sock.fd)) return rc;
    const size_t n3 = (size_t)3 * na;
    std::vector<double> r(n3), dvdr(n3), dvdr2(n3), frc(n3);
    double cell[9], icell[9], box[3] = {0, 0, 0}, v = 0.0, vew, voo, vint, vir[9], send_vir[9];
    bool hasdata = false;
    for (;;) {
        char h[MSGLEN + 1] = {0};
        // the stream ended: the server is done (the reference leaves its loop, src/ipi_driver.f90:118-122)
        if (!read_all(sock.fd, h, MSGLEN)) return 0;
        for (int i = MSGLEN - 1; i >= 0 && (h[i] == ' ' || h[i] == '\0'); i--) h[i] = '\0';
        const std::string msg(h);
        if (msg == "STATUS") {
            if (!write_header(sock.fd, hasdata ? "HAVEDATA" : "READY")) return RPMD_ERR_SOCKET;
        } else if (msg == "POSDATA") {
            int32_t nat = 0;
            if (!read_all(sock.fd, cell, sizeof cell) || !read_all(sock.fd, icell, sizeof icell) ||
                !read_all(sock.fd, &nat, sizeof nat))
                return RPMD_ERR_SOCKET;
            if (nat != na) return RPMD_ERR_IPI_NAT;
            if (!read_all(sock.fd, r.data(), n3 * sizeof(double))) return RPMD_ERR_SOCKET;
            // the orthorhombic box: the cell's diagonal (src/ipi_driver.f90:150-153)
            box[0] = cell[0];
            box[1] = cell[4];
            box[2] = cell[8];
            if (int rc = rpmd_full_forces(ctx, r.data(), na, 1, &v, &vew, &voo, &vint, vir, z, box, dvdr.data(),
                                          dvdr2.data()))
                return rc;
            for (size_t i = 0; i < n3; i++) frc[i] = -dvdr[i] - dvdr2[i];
            // minus the transposed virial, in the reference's column order (src/ipi_driver.f90:182-183)
            for (int i = 0; i < 3; i++)
                for (int j = 0; j < 3; j++) send_vir[j * 3 + i] = -vir[i * 3 + j];
            hasdata = true;
        } else if (msg == "GETFORCE") {
            const int32_t nat = na, extra = 0;
            if (!write_header(sock.fd, "FORCEREADY") || !write_all(sock.fd, &v, sizeof v) ||
                !write_all(sock.fd, &nat, sizeof nat) || !write_all(sock.fd, frc.data(), n3 * sizeof(double)) ||
                !write_all(sock.fd, send_vir, sizeof send_vir) || !write_all(sock.fd, &extra, sizeof extra))
                return RPMD_ERR_SOCKET;
            hasdata = false;
            if (served) (*served)++;
        } else {
            // EXIT or anything else: the reference leaves its loop (src/ipi_driver.f90:237-240)
            return 0;
        }
    }
}
