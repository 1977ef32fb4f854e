/*
 * Internal entry points shared by the C host files of libmyers_amd.so
 * (not exported in the public headers).
 */
#ifndef MYERS_INTERNAL_H
#define MYERS_INTERNAL_H

/* 1 if every operation enqueued on the buffer's stream has finished, 0 if
 * work is pending, negative on a HIP error. */
int myers_buffer_ready(void *myersBuffer);

#endif
