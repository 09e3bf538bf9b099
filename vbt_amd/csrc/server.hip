// server.hip placeholder
