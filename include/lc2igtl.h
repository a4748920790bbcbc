/*
 * lc2igtl.h -- OpenIGTLink TRANSFORM streaming of the registration pose.
 *
 * The reference's Interpolator3D opens an OpenIGTLink server on TCP port 23333, waits for
 * one client (interpolator.cu:359-388) and, at every GenerateMappers, sends the current
 * transform -- GetMyselfInversed of VirtualToMoving, a 4x4 -- as a TRANSFORM message with
 * header version 2 and device name "RegistrationServer" (interpolator.cu:545-573). The
 * OpenIGTLink library is not available to this build; these entry points write the
 * protocol's bytes themselves:
 *   header (58 bytes, big-endian): version u16, type char[12] "TRANSFORM", device name
 *     char[20], time stamp u64 (seconds << 32 | fraction of 2^32), body size u64, CRC-64
 *     (ECMA-182 polynomial 0x42F0E1EBA9EA3693, initial value 0, no reflection) of the body;
 *   body of header version 2: extended header (size u16 = 12, metadata header size u16,
 *     metadata size u32, message id u32), the content -- 12 float32 R11 R21 R31 R12 R22 R32
 *     R13 R23 R33 TX TY TZ (rotation column by column, then translation) -- and an empty
 *     metadata header (index count u16 = 0); version 1 carries the content only.
 * Statuses: 0 ok, negative on error (message in lc2igtl_last_error()).
 */
#ifndef LC2IGTL_H
#define LC2IGTL_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lc2igtl_server lc2igtl_server;

const char* lc2igtl_last_error(void);

/* Bytes of a TRANSFORM message for the row-major 4x4 `m16` (its top 3x4 is sent).
 * header_version 1 or 2; out may be NULL to query the size into *len. */
int lc2igtl_pack_transform(const char* device_name, const float m16[16], double timestamp_s,
                           int header_version, unsigned message_id, unsigned char* out,
                           size_t cap, size_t* len);

/* A server listening on 127.0.0.1-or-any:port (0 = an ephemeral port; the reference uses
 * 23333). lc2igtl_port reports the port it listens on. */
int lc2igtl_server_create(int port, int loopback_only, lc2igtl_server** out);
int lc2igtl_port(lc2igtl_server* s);
/* Wait for the client (timeout_ms < 0: no limit, as the reference's loop); 1 = connected,
 * 0 = time-out. */
int lc2igtl_wait_client(lc2igtl_server* s, double timeout_ms);
/* Send one TRANSFORM (header version 2, message ids counting from 0) to the connected
 * client; the time stamp is the wall clock. */
int lc2igtl_send_transform(lc2igtl_server* s, const char* device_name, const float m16[16]);
/* Messages sent so far. */
long long lc2igtl_sent(lc2igtl_server* s);
void lc2igtl_server_destroy(lc2igtl_server* s);

#ifdef __cplusplus
}
#endif
#endif
